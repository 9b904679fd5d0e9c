// io.cpp — projection-file reader/writer and image writers of the reference's formats.
//
// Projection files (count_histories_v0, DS.cu:3578-3678; read_data_chunk_v0 :4361-4467;
// SupplementalDocumentation/ProposedDataFormat_Version1-1.txt):
//   "PCTD" | "PCTN" (4 bytes), version (int), N (int), projection angle (float, degrees),
//   beam energy (float), generation date (int), preprocessing date (int),
//   three length-prefixed strings (phantom name, data source, prepared by),
//   then 13 arrays of N floats: t0 t1 t2 t3 v0 v1 v2 v3 u0 u1 u2 u3 WEPL,
//   and for version 1 (and continuous-angle PCTN data) N per-history angles (float, degrees).
// Images: array_2_disk text layout (DS.cu:10743-10787) and write_PNG's .dat (DS.cu:9939-9973), plus the PNG
// that write_PNG makes from the .dat with ImageMagick (DS.cu:9969-9971).
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
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
namespace {
// Joins every thread it holds when it goes out of scope, on error paths too.
struct ThreadGroup {
    std::vector<std::thread> t;
    ~ThreadGroup()
    {
        for (auto& th : t)
            if (th.joinable()) th.join();
    }
};
}  // namespace

extern "C" int pct_write_image_text(const char* path, const float* image, int columns, int rows, int slices)
{
    if (!path || !image || columns <= 0 || rows <= 0 || slices <= 0) return PCT_E_ARG;
    try {
        std::ofstream out(path);
        if (!out) return PCT_E_IO;
        // Each slice is formatted by its own default-format stream (the format of `out`: precision 6,
        // the global locale), up to 16 slices at a time on separate threads, and the slices are written
        // in order.  '\n' writes the byte std::endl writes, without a flush per line.
        const int64_t plane = (int64_t)columns * rows;
        std::atomic<bool> failed{false};
        auto format_slice = [&](int z, std::string& dst) {
            try {
                std::ostringstream ss;
                const float* p = image + (int64_t)z * plane;
                for (int y = 0; y < rows; y++) {
                    for (int x = 0; x < columns; x++) ss << p[x + (int64_t)y * columns] << " ";
                    ss << '\n';
                }
                dst = ss.str();
            } catch (...) {
                failed = true;
            }
        };
        unsigned hw = std::thread::hardware_concurrency();
        const int width = (int)std::min(16u, hw > 0 ? hw : 1u);
        std::vector<std::string> text(width);
        for (int z0 = 0; z0 < slices; z0 += width) {
            const int nb = std::min(width, slices - z0);
            {
                ThreadGroup pool;
                pool.t.reserve(nb);
                for (int k = 1; k < nb; k++) {
                    try {
                        pool.t.emplace_back(format_slice, z0 + k, std::ref(text[k]));
                    } catch (...) {
                        format_slice(z0 + k, text[k]);  // no thread available: format here
                    }
                }
                format_slice(z0, text[0]);
            }  // joins the slice threads
            if (failed) return PCT_E_NOMEM;
            for (int k = 0; k < nb; k++) out.write(text[k].data(), (std::streamsize)text[k].size());
            if (!out) return PCT_E_IO;
        }
        out.flush();
        return out ? PCT_OK : PCT_E_IO;
    } catch (const std::bad_alloc&) {
        return PCT_E_NOMEM;
    } catch (...) {
        return PCT_E_IO;
    }
}

// write_PNG's binary file: float32 value/2, x fastest, then y, then z.
extern "C" int pct_write_image_dat(const char* path, const float* image, int columns, int rows, int slices)
{
    if (!path || !image || columns <= 0 || rows <= 0 || slices <= 0) return PCT_E_ARG;
    FILE* f = std::fopen(path, "wb");
    if (!f) return PCT_E_IO;
    FileCloser fc{f};
    const int64_t nv = (int64_t)columns * rows * slices;
    // in pieces of 1M values: bounded memory for any image size
    constexpr int64_t kPiece = 1 << 20;
    std::vector<float> buf;
    try {
        buf.resize((size_t)std::min(nv, kPiece));
    } catch (...) {
        return PCT_E_NOMEM;
    }
    for (int64_t i0 = 0; i0 < nv; i0 += kPiece) {
        const int64_t m = std::min(kPiece, nv - i0);
        for (int64_t i = 0; i < m; i++) buf[i] = image[i0 + i] / 2;
        if (std::fwrite(buf.data(), sizeof(float), (size_t)m, f) != (size_t)m) return PCT_E_IO;
    }
    return PCT_OK;
}

// write_PNG's PNG (DS.cu:9969-9971): ImageMagick `convert -define quantum:format=floating-point -depth 32
// -size COLUMNS x ROWS*SLICES gray:file.dat file.png` reads the .dat floats (value/2) as gray levels with
// 1.0 = white and writes a 16-bit grayscale PNG.  Here: the same COLUMNS x ROWS*SLICES image, sample =
// round(clamp(value/2, 0, 1) * 65535), written directly (filter 0, zlib stream of stored deflate blocks).
namespace {
struct CrcTable {
    uint32_t v[256];
    CrcTable()
    {
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t c = i;
            for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            v[i] = c;
        }
    }
};
uint32_t crc32_update(uint32_t c, const uint8_t* p, size_t n)
{
    static const CrcTable table;  // initialised once, thread-safe (function-local static)
    for (size_t i = 0; i < n; i++) c = table.v[(c ^ p[i]) & 0xFF] ^ (c >> 8);
    return c;
}
void put_be32(std::vector<uint8_t>* o, uint32_t v)
{
    for (int s = 24; s >= 0; s -= 8) o->push_back((uint8_t)(v >> s));
}
// One chunk of `len` bytes at `data` (a PNG chunk holds at most 2^31 - 1 bytes; callers split).
bool write_chunk(FILE* f, const char type[4], const uint8_t* data, size_t len)
{
    std::vector<uint8_t> head;
    put_be32(&head, (uint32_t)len);
    head.insert(head.end(), type, type + 4);
    uint32_t crc = crc32_update(0xFFFFFFFFu, (const uint8_t*)type, 4);
    crc = crc32_update(crc, data, len) ^ 0xFFFFFFFFu;
    std::vector<uint8_t> tail;
    put_be32(&tail, crc);
    return std::fwrite(head.data(), 1, head.size(), f) == head.size() &&
           (len == 0 || std::fwrite(data, 1, len, f) == len) &&
           std::fwrite(tail.data(), 1, tail.size(), f) == tail.size();
}
}  // namespace

extern "C" int pct_write_image_png(const char* path, const float* image, int columns, int rows, int slices)
{
    if (!path || !image || columns <= 0 || rows <= 0 || slices <= 0) return PCT_E_ARG;
    const int64_t height = (int64_t)rows * slices;
    if (height > 0x7FFFFFFF || columns > 0x3FFFFFFF) return PCT_E_LIMIT;
    try {
        // raw scanlines: filter byte 0, then big-endian 16-bit samples
        const size_t line = 1 + 2 * (size_t)columns;
        std::vector<uint8_t> raw(line * (size_t)height);
        for (int64_t y = 0; y < height; y++) {
            uint8_t* p = raw.data() + (size_t)y * line;
            *p++ = 0;
            for (int x = 0; x < columns; x++) {
                float v = image[x + y * columns] / 2;
                v = v > 0.0f ? (v < 1.0f ? v : 1.0f) : 0.0f;  // NaN -> 0
                const uint16_t q = (uint16_t)(v * 65535.0f + 0.5f);
                *p++ = (uint8_t)(q >> 8);
                *p++ = (uint8_t)(q & 0xFF);
            }
        }
        // zlib stream: header, stored deflate blocks of at most 65535 bytes, Adler-32
        std::vector<uint8_t> z;
        z.reserve(raw.size() + raw.size() / 65535 * 5 + 16);
        z.push_back(0x78);
        z.push_back(0x01);
        size_t pos = 0;
        do {
            const size_t len = std::min<size_t>(65535, raw.size() - pos);
            const bool last = pos + len == raw.size();
            z.push_back(last ? 1 : 0);
            z.push_back((uint8_t)(len & 0xFF));
            z.push_back((uint8_t)(len >> 8));
            z.push_back((uint8_t)(~len & 0xFF));
            z.push_back((uint8_t)((~len >> 8) & 0xFF));
            z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + len);
            pos += len;
        } while (pos < raw.size());
        uint32_t a = 1, b = 0;
        for (uint8_t c : raw) {
            a = (a + c) % 65521u;
            b = (b + a) % 65521u;
        }
        put_be32(&z, (b << 16) | a);
        std::vector<uint8_t> ihdr;
        put_be32(&ihdr, (uint32_t)columns);
        put_be32(&ihdr, (uint32_t)height);
        ihdr.push_back(16);  // bit depth
        ihdr.push_back(0);   // grayscale
        ihdr.push_back(0);   // deflate
        ihdr.push_back(0);   // filter method 0
        ihdr.push_back(0);   // no interlace
        FILE* f = std::fopen(path, "wb");
        if (!f) return PCT_E_IO;
        FileCloser fc{f};
        static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
        if (std::fwrite(sig, 1, 8, f) != 8) return PCT_E_IO;
        if (!write_chunk(f, "IHDR", ihdr.data(), ihdr.size())) return PCT_E_IO;
        // the zlib stream as consecutive IDAT chunks of at most 1 MiB (decoders concatenate them)
        constexpr size_t kIdat = 1u << 20;
        for (size_t o = 0; o < z.size(); o += kIdat)
            if (!write_chunk(f, "IDAT", z.data() + o, std::min(kIdat, z.size() - o))) return PCT_E_IO;
        if (!write_chunk(f, "IEND", nullptr, 0)) return PCT_E_IO;
        return PCT_OK;
    } catch (const std::bad_alloc&) {
        return PCT_E_NOMEM;
    }
}
