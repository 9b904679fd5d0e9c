// io.cpp — projection-file reader/writer and image writers of the reference's formats.
//
// Projection files (count_histories_v0, DS.cu:3578-3678; read_data_chunk_v0 :4361-4467;
// SupplementalDocumentation/ProposedDataFormat_Version1-1.txt):
//   "PCTD" | "PCTN" (4 bytes), version (int), N (int), projection angle (float, degrees),
//   beam energy (float), generation date (int), preprocessing date (int),
//   three length-prefixed strings (phantom name, data source, prepared by),
//   then 13 arrays of N floats: t0 t1 t2 t3 v0 v1 v2 v3 u0 u1 u2 u3 WEPL,
//   and for version 1 (and continuous-angle PCTN data) N per-history angles (float, degrees).
// Images: array_2_disk text layout (DS.cu:10743-10787) and write_PNG's .dat (DS.cu:9939-9973).
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/pct.h"

namespace {
struct FileCloser {
    FILE* f;
    ~FileCloser()
    {
        if (f) std::fclose(f);
    }
};
bool read_exact(FILE* f, void* dst, size_t bytes) { return std::fread(dst, 1, bytes, f) == bytes; }
}  // namespace

extern "C" int pct_read_projection(const char* path, int64_t* n, float* header_angle_deg, float* const t[4],
                                   float* const v[4], float* const u[4], float* wepl, float* angles)
{
    if (!path || !n) return PCT_E_ARG;
    FILE* f = std::fopen(path, "rb");
    if (!f) return PCT_E_IO;
    FileCloser fc{f};
    char magic[5] = {0};
    if (!read_exact(f, magic, 4)) return PCT_E_IO;
    bool continuous = std::strcmp(magic, "PCTN") == 0;
    if (!continuous && std::strcmp(magic, "PCTD") != 0) return PCT_E_IO;  // DS.cu:3611-3618
    int32_t version = 0, count = 0, gen_date = 0, pre_date = 0;
    float angle = 0.0f, energy = 0.0f;
    if (!read_exact(f, &version, 4)) return PCT_E_IO;
    if (version != 0 && version != 1) return PCT_E_IO;  // DS.cu:3671-3675
    if (!read_exact(f, &count, 4) || !read_exact(f, &angle, 4) || !read_exact(f, &energy, 4) ||
        !read_exact(f, &gen_date, 4) || !read_exact(f, &pre_date, 4))
        return PCT_E_IO;
    if (count < 0) return PCT_E_IO;
    for (int k = 0; k < 3; k++) {
        int32_t len = 0;
        if (!read_exact(f, &len, 4) || len < 0) return PCT_E_IO;
        if (std::fseek(f, len, SEEK_CUR) != 0) return PCT_E_IO;
    }
    *n = count;
    if (header_angle_deg) *header_angle_deg = angle;
    if (!t || !v || !u || !wepl) return PCT_OK;  // header only
    float* const dst[13] = {t[0], t[1], t[2], t[3], v[0], v[1], v[2], v[3], u[0], u[1], u[2], u[3], wepl};
    for (int k = 0; k < 13; k++)
        if (!dst[k] || !read_exact(f, dst[k], (size_t)count * sizeof(float))) return PCT_E_IO;
    bool has_angles = version == 1 || continuous;
    if (angles) {
        if (has_angles) {
            if (!read_exact(f, angles, (size_t)count * sizeof(float))) return PCT_E_IO;
        } else {
            // integer gantry angle of the file (gantry_angle_h = int(projection_angle), DS.cu:4432)
            for (int32_t i = 0; i < count; i++) angles[i] = (float)(int)angle;
        }
    }
    return PCT_OK;
}

extern "C" int pct_write_projection(const char* path, int64_t n, float angle_deg, float energy_mev,
                                    const float* const t[4], const float* const v[4], const float* const u[4],
                                    const float* wepl, const float* angles)
{
    if (!path || n < 0 || n > 0x7fffffff || !t || !v || !u || !wepl) return PCT_E_ARG;
    FILE* f = std::fopen(path, "wb");
    if (!f) return PCT_E_IO;
    FileCloser fc{f};
    const char* magic = angles ? "PCTN" : "PCTD";
    int32_t version = angles ? 1 : 0, count = (int32_t)n, zero = 0;
    bool ok = std::fwrite(magic, 1, 4, f) == 4;
    ok = ok && std::fwrite(&version, 4, 1, f) == 1 && std::fwrite(&count, 4, 1, f) == 1;
    ok = ok && std::fwrite(&angle_deg, 4, 1, f) == 1 && std::fwrite(&energy_mev, 4, 1, f) == 1;
    ok = ok && std::fwrite(&zero, 4, 1, f) == 1 && std::fwrite(&zero, 4, 1, f) == 1;
    const char* strs[3] = {"synthetic phantom", "pct_reconstruction_amd", "pct_reconstruction_amd"};
    for (const char* s : strs) {
        int32_t len = (int32_t)std::strlen(s);
        ok = ok && std::fwrite(&len, 4, 1, f) == 1 && std::fwrite(s, 1, len, f) == (size_t)len;
    }
    const float* const src[13] = {t[0], t[1], t[2], t[3], v[0], v[1], v[2], v[3], u[0], u[1], u[2], u[3], wepl};
    for (int k = 0; k < 13 && ok; k++) ok = std::fwrite(src[k], sizeof(float), (size_t)n, f) == (size_t)n;
    if (angles && ok) ok = std::fwrite(angles, sizeof(float), (size_t)n, f) == (size_t)n;
    return ok ? PCT_OK : PCT_E_IO;
}

// array_2_disk: one line per (y, z) row with `value ' '` for each x, default ostream formatting.
extern "C" int pct_write_image_text(const char* path, const float* image, int columns, int rows, int slices)
{
    if (!path || !image || columns <= 0 || rows <= 0 || slices <= 0) return PCT_E_ARG;
    std::ofstream out(path);
    if (!out) return PCT_E_IO;
    for (int z = 0; z < slices; z++)
        for (int y = 0; y < rows; y++) {
            for (int x = 0; x < columns; x++) out << image[x + (int64_t)y * columns + (int64_t)z * columns * rows] << " ";
            out << std::endl;
        }
    return out ? PCT_OK : PCT_E_IO;
}

// write_PNG's binary file: float32 value/2, x fastest, then y, then z.
extern "C" int pct_write_image_dat(const char* path, const float* image, int columns, int rows, int slices)
{
    if (!path || !image || columns <= 0 || rows <= 0 || slices <= 0) return PCT_E_ARG;
    FILE* f = std::fopen(path, "wb");
    if (!f) return PCT_E_IO;
    FileCloser fc{f};
    const int64_t nv = (int64_t)columns * rows * slices;
    std::vector<float> buf(nv);
    for (int64_t i = 0; i < nv; i++) buf[i] = image[i] / 2;
    return std::fwrite(buf.data(), sizeof(float), nv, f) == (size_t)nv ? PCT_OK : PCT_E_IO;
}
