// checkpoint.cpp -- the reference's restart file, format v1
// (Core/Checkpoint.hpp:37-234), read and written field by field.
//
// Layout (little-endian, no padding beyond what the reference's structs
// carry):
//   header   56 B   magic u64 "ALEBGK01", version u32, dim u32, N u64,
//                   gStride u64, step u64, t f64, nBodies u32, pad u32
//   N x 200 B       particle records: x y z ux uy uz T rho p meshV(3) wn(3)
//                   wU(3) wT vlo vhi dv_local (22 f64), Nv_local dim bodyId
//                   voxel (4 i32), boundary validg moodFlag pad (4 u8), 4 B pad
//   N x gStride f64 g, particle by particle
//   nBodies x 232 B body records: cx cy cz velx vely velz angularVel omx omy
//                   omz forcex forcey forcez torque tqx tqy tqz ox oy (19 f64),
//                   Rm[9] f64, dim i32, rotates u8, 3 B pad
// Each field goes to its byte offset explicitly, so the file does not depend
// on this compiler's struct layout. As the reference does, a write goes to
// path.tmp and is renamed over path; a read refuses a file whose N, gStride
// or body count differ from what the caller expects.
#include "alebgk_host.h"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr uint64_t kMagic = 0x414c4542474b3031ULL;   // "ALEBGK01"
constexpr uint32_t kVersion = 1;
constexpr size_t kHeader = 56, kPRec = 200, kBRec = 232;

template <typename T>
void put(unsigned char *b, size_t off, T v) { std::memcpy(b + off, &v, sizeof v); }
template <typename T>
T get(const unsigned char *b, size_t off) { T v; std::memcpy(&v, b + off, sizeof v); return v; }

struct File
{
    FILE *f = nullptr;
    explicit File(FILE *h) : f(h) {}
    ~File() { if (f) std::fclose(f); }
};

void encode_particle(unsigned char *r, const alh_ckpt_fields *in, size_t i)
{
    std::memset(r, 0, kPRec);
    const double d[22] = {in->x[i], in->y[i], in->z[i],
                          in->mom[6 * i + 1], in->mom[6 * i + 2], in->mom[6 * i + 3],
                          in->mom[6 * i + 4], in->mom[6 * i + 0], in->mom[6 * i + 5],
                          in->mesh[3 * i], in->mesh[3 * i + 1], in->mesh[3 * i + 2],
                          in->wall[7 * i], in->wall[7 * i + 1], in->wall[7 * i + 2],
                          in->wall[7 * i + 3], in->wall[7 * i + 4], in->wall[7 * i + 5],
                          in->wall[7 * i + 6],
                          in->grid[3 * i], in->grid[3 * i + 1], in->grid[3 * i + 2]};
    for (int k = 0; k < 22; ++k) put<double>(r, 8 * k, d[k]);
    for (int k = 0; k < 4; ++k) put<int32_t>(r, 176 + 4 * k, in->ints[4 * i + k]);
    for (int k = 0; k < 3; ++k) r[192 + k] = in->flags[3 * i + k] ? 1 : 0;
}

void decode_particle(const unsigned char *r, alh_ckpt_fields *out, size_t i)
{
    double d[22];
    for (int k = 0; k < 22; ++k) d[k] = get<double>(r, 8 * k);
    out->x[i] = d[0]; out->y[i] = d[1]; out->z[i] = d[2];
    out->mom[6 * i + 1] = d[3]; out->mom[6 * i + 2] = d[4]; out->mom[6 * i + 3] = d[5];
    out->mom[6 * i + 4] = d[6]; out->mom[6 * i + 0] = d[7]; out->mom[6 * i + 5] = d[8];
    for (int k = 0; k < 3; ++k) out->mesh[3 * i + k] = d[9 + k];
    for (int k = 0; k < 7; ++k) out->wall[7 * i + k] = d[12 + k];
    for (int k = 0; k < 3; ++k) out->grid[3 * i + k] = d[19 + k];
    for (int k = 0; k < 4; ++k) out->ints[4 * i + k] = get<int32_t>(r, 176 + 4 * k);
    for (int k = 0; k < 3; ++k) out->flags[3 * i + k] = r[192 + k] ? 1 : 0;
}

void encode_body(unsigned char *r, const double *b)
{
    std::memset(r, 0, kBRec);
    for (int k = 0; k < 28; ++k) put<double>(r, 8 * k, b[k]);
    put<int32_t>(r, 224, (int32_t)b[28]);
    r[228] = b[29] != 0.0 ? 1 : 0;
}

void decode_body(const unsigned char *r, double *b)
{
    for (int k = 0; k < 28; ++k) b[k] = get<double>(r, 8 * k);
    b[28] = (double)get<int32_t>(r, 224);
    b[29] = r[228] ? 1.0 : 0.0;
}

int read_header(FILE *f, alh_ckpt_info *info)
{
    unsigned char h[kHeader];
    if (std::fread(h, 1, kHeader, f) != kHeader) return AB_ERR_ARG;
    if (get<uint64_t>(h, 0) != kMagic || get<uint32_t>(h, 8) != kVersion) return AB_ERR_ARG;
    info->dim = (int)get<uint32_t>(h, 12);
    info->n = (long long)get<uint64_t>(h, 16);
    info->gstride = (long long)get<uint64_t>(h, 24);
    info->step = (long long)get<uint64_t>(h, 32);
    info->t = get<double>(h, 40);
    info->nbodies = (int)get<uint32_t>(h, 48);
    return AB_OK;
}

}  // namespace

extern "C" {

int alh_ckpt_info_read(const char *path, alh_ckpt_info *info)
{
    if (!path || !info) return AB_ERR_ARG;
    File f(std::fopen(path, "rb"));
    if (!f.f) return AB_ERR_ARG;
    return read_header(f.f, info);
}

int alh_ckpt_read(const char *path, const alh_ckpt_info *expect, alh_ckpt_fields *out)
{
    if (!path || !expect || !out || !out->x || !out->y || !out->z || !out->mom || !out->mesh ||
        !out->wall || !out->grid || !out->ints || !out->flags || !out->g ||
        (expect->nbodies > 0 && !out->bodies))
        return AB_ERR_ARG;
    File f(std::fopen(path, "rb"));
    if (!f.f) return AB_ERR_ARG;
    alh_ckpt_info h{};
    int rc = read_header(f.f, &h);
    if (rc) return rc;
    // a different resolution or body count is refused (Checkpoint.hpp:173-186)
    if (h.n != expect->n || h.gstride != expect->gstride || h.nbodies != expect->nbodies ||
        h.n <= 0)
        return AB_ERR_UNSUPPORTED;
    const size_t N = (size_t)h.n, gs = (size_t)h.gstride;
    std::vector<unsigned char> rec(N * kPRec);
    if (std::fread(rec.data(), 1, rec.size(), f.f) != rec.size()) return AB_ERR_ARG;
    for (size_t i = 0; i < N; ++i) decode_particle(rec.data() + i * kPRec, out, i);
    if (std::fread(out->g, sizeof(double), N * gs, f.f) != N * gs) return AB_ERR_ARG;
    if (h.nbodies > 0)
    {
        std::vector<unsigned char> br((size_t)h.nbodies * kBRec);
        if (std::fread(br.data(), 1, br.size(), f.f) != br.size()) return AB_ERR_ARG;
        for (int b = 0; b < h.nbodies; ++b) decode_body(br.data() + (size_t)b * kBRec, out->bodies + 30 * b);
    }
    return AB_OK;
}

int alh_ckpt_write(const char *path, const alh_ckpt_info *info, const alh_ckpt_fields *in)
{
    if (!path || !info || !in || info->n <= 0 || info->gstride <= 0 || info->nbodies < 0 ||
        !in->x || !in->y || !in->z || !in->mom || !in->mesh || !in->wall || !in->grid ||
        !in->ints || !in->flags || !in->g || (info->nbodies > 0 && !in->bodies))
        return AB_ERR_ARG;
    const std::string tmp = std::string(path) + ".tmp";
    const size_t N = (size_t)info->n, gs = (size_t)info->gstride;
    {
        File f(std::fopen(tmp.c_str(), "wb"));
        if (!f.f) return AB_ERR_ARG;
        unsigned char h[kHeader];
        std::memset(h, 0, sizeof h);
        put<uint64_t>(h, 0, kMagic);
        put<uint32_t>(h, 8, kVersion);
        put<uint32_t>(h, 12, (uint32_t)info->dim);
        put<uint64_t>(h, 16, (uint64_t)N);
        put<uint64_t>(h, 24, (uint64_t)gs);
        put<uint64_t>(h, 32, (uint64_t)info->step);
        put<double>(h, 40, info->t);
        put<uint32_t>(h, 48, (uint32_t)info->nbodies);
        bool ok = std::fwrite(h, 1, kHeader, f.f) == kHeader;
        std::vector<unsigned char> rec(N * kPRec);
        for (size_t i = 0; i < N; ++i) encode_particle(rec.data() + i * kPRec, in, i);
        ok = ok && std::fwrite(rec.data(), 1, rec.size(), f.f) == rec.size();
        ok = ok && std::fwrite(in->g, sizeof(double), N * gs, f.f) == N * gs;
        std::vector<unsigned char> br((size_t)info->nbodies * kBRec);
        for (int b = 0; b < info->nbodies; ++b) encode_body(br.data() + (size_t)b * kBRec, in->bodies + 30 * b);
        if (!br.empty()) ok = ok && std::fwrite(br.data(), 1, br.size(), f.f) == br.size();
        ok = ok && std::fflush(f.f) == 0;
        if (!ok) { std::fclose(f.f); f.f = nullptr; std::remove(tmp.c_str()); return AB_ERR_ARG; }
    }
    std::remove(path);
    if (std::rename(tmp.c_str(), path) != 0) return AB_ERR_ARG;
    return AB_OK;
}

}  // extern "C"
