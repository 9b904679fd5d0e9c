// preprocessed.hip — export and import of the preprocessed data (path lists and what DROP needs).
//
// The reference's variant program (src/pCT_Reconstruction.cu, "V.cu") persists the MLP paths and
// their per-history data so that reconstruction can start from them:
//   accumulate_and_export_preprocessed_data V.cu:6613-6658, export_hull :6540, export_x_0 :6551,
//   export_WEPL :6562, export_histories :6571, export_voxels_per_path :6594,
//   export_avg_chord_lengths :6603; import_hull :6659, import_x_0 :6672, import_WEPL :6681,
//   import_histories :6693, import_voxels_per_path :6728, import_avg_chord_lengths :6740,
//   import_MLP :6752, init_import_preprocessed_data :6775-6861.
// File names: MLP_BASENAME etc. (include/Constants.h:84-88) with the write-mode extensions
// (include/Configurations.h:1802-1822, include/Constants.h:98-106).
//
// Layout written here: every file starts with the history count as uint32 (the form all the
// importers read); MLP.bin then holds the concatenated paths as uint32 voxel indices.  The reference's
// own exporter writes MLP.bin and WEPL.bin with a "%u\n" text header (V.cu:6628-6629), which its
// importer then misreads; the importer here accepts either header, telling them apart by the file
// size.  The lists move between device and file in chunks of at most kChunk entries.
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"

namespace {

constexpr int64_t kChunk = 1ll << 26;  // list entries per device<->host transfer (256 MB)

struct File {
    FILE* f = nullptr;
    ~File()
    {
        if (f) std::fclose(f);
    }
};

std::string join(const char* dir, const char* name)
{
    std::string p(dir);
    if (!p.empty() && p.back() != '/') p += '/';
    return p + name;
}

int io_fail(pct_ctx* ctx, const std::string& what)
{
    ctx->err = what;
    return PCT_E_IO;
}

template <class T>
int write_counted(pct_ctx* ctx, const std::string& path, uint32_t n, const std::vector<T>& v)
{
    File out;
    out.f = std::fopen(path.c_str(), "wb");
    if (!out.f) return io_fail(ctx, "cannot write " + path + ": " + std::strerror(errno));
    if (std::fwrite(&n, sizeof(uint32_t), 1, out.f) != 1 || std::fwrite(v.data(), sizeof(T), v.size(), out.f) != v.size())
        return io_fail(ctx, "short write to " + path);
    return PCT_OK;
}

template <class T>
int copy_to_host(pct_ctx* ctx, const T* src, int64_t n, std::vector<T>* dst)
{
    dst->resize((size_t)n);
    if (n > 0) PCT_CHECK_HIP(ctx, hipMemcpy(dst->data(), src, n * sizeof(T), hipMemcpyDeviceToHost));
    return PCT_OK;
}

long long file_size(FILE* f)
{
    if (std::fseek(f, 0, SEEK_END) != 0) return -1;
    long long s = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    return s;
}

// Header of MLP.bin / WEPL.bin: the uint32 count, or the exporter's "%u\n" text.  `payload` is the
// number of bytes the data must occupy after the header for n histories.  Returns the header length.
long long read_header(FILE* f, uint32_t n, long long payload, long long size)
{
    unsigned char b[12] = {};
    const size_t got = std::fread(b, 1, sizeof(b), f);
    if (got >= 4) {
        uint32_t v;
        std::memcpy(&v, b, 4);
        if (v == n && size == 4 + payload) return 4;
    }
    size_t k = 0;
    unsigned long long t = 0;
    while (k < got && k < 11 && b[k] >= '0' && b[k] <= '9') t = t * 10 + (b[k++] - '0');
    if (k > 0 && k < got && b[k] == '\n' && t == n && size == (long long)(k + 1) + payload) return (long long)(k + 1);
    return -1;
}

// The longest path the library's own MLP walk can produce; imported paths must not exceed it.
int64_t max_path(const pct_params& P) { return 4ll * (P.columns + P.rows + P.slices) + 16; }

}  // namespace

extern "C" int pct_export_preprocessed(pct_ctx* ctx, const char* dir, int flags)
{
    if (!ctx || !dir || (flags & ~PCT_EXPORT_CHORD_F64)) return PCT_E_ARG;
    if (ctx->stage < ST_LISTS) {
        ctx->err = "pct_export_preprocessed: the MLP lists have not been built";
        return PCT_E_STATE;
    }
    const pct_params& P = ctx->P;
    const int64_t n = ctx->n_recon;
    if (n > (int64_t)UINT32_MAX) return PCT_E_LIMIT;
    const uint32_t n32 = (uint32_t)n;
    PCT_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));

    // per-history arrays
    std::vector<int64_t> off;
    std::vector<uint32_t> cnt, first;
    std::vector<double> chord;
    std::vector<float> wepl, hf[10], gantry;
    std::vector<int32_t> bin;
    PCT_TRY(copy_to_host(ctx, ctx->list_off, n + 1, &off));
    PCT_TRY(copy_to_host(ctx, ctx->list_cnt, n, &cnt));
    PCT_TRY(copy_to_host(ctx, ctx->chord, n, &chord));
    PCT_TRY(copy_to_host(ctx, ctx->wepl, n, &wepl));
    PCT_TRY(copy_to_host(ctx, ctx->first_voxel, n, &first));
    PCT_TRY(copy_to_host(ctx, ctx->bin_num, n, &bin));
    PCT_TRY(copy_to_host(ctx, ctx->gantry, n, &gantry));
    for (int k = 0; k < 10; k++) PCT_TRY(copy_to_host(ctx, ctx->h[k], n, &hf[k]));

    // MLP.bin: the paths in history order, streamed in chunks of whole histories
    {
        const std::string path = join(dir, "MLP.bin");
        File out;
        out.f = std::fopen(path.c_str(), "wb");
        if (!out.f) return io_fail(ctx, "cannot write " + path + ": " + std::strerror(errno));
        if (std::fwrite(&n32, sizeof(uint32_t), 1, out.f) != 1) return io_fail(ctx, "short write to " + path);
        std::vector<uint32_t> buf;
        int64_t i = 0;
        while (i < n) {
            int64_t j = i + 1;
            while (j < n && off[j + 1] - off[i] <= kChunk) j++;
            const int64_t span = off[j] - off[i];
            buf.resize((size_t)span);
            if (span > 0)
                PCT_CHECK_HIP(ctx, hipMemcpy(buf.data(), ctx->list_vox + off[i], span * sizeof(uint32_t),
                                             hipMemcpyDeviceToHost));
            for (int64_t h = i; h < j; h++)
                if (cnt[h] && std::fwrite(buf.data() + (off[h] - off[i]), sizeof(uint32_t), cnt[h], out.f) != cnt[h])
                    return io_fail(ctx, "short write to " + path);
            i = j;
        }
    }
    PCT_TRY(write_counted(ctx, join(dir, "voxels_per_path.bin"), n32, cnt));
    std::vector<float> chord32(chord.begin(), chord.end());
    PCT_TRY(write_counted(ctx, join(dir, "avg_chord_lengths.bin"), n32, chord32));
    if (flags & PCT_EXPORT_CHORD_F64) PCT_TRY(write_counted(ctx, join(dir, "avg_chord_lengths_f64.bin"), n32, chord));
    PCT_TRY(write_counted(ctx, join(dir, "WEPL.bin"), n32, wepl));

    // histories.bin (export_histories :6571-6592): entry voxel x, y, z; 10 coordinate/angle arrays;
    // bin number; gantry angle
    {
        const std::string path = join(dir, "histories.bin");
        File out;
        out.f = std::fopen(path.c_str(), "wb");
        if (!out.f) return io_fail(ctx, "cannot write " + path + ": " + std::strerror(errno));
        bool ok = std::fwrite(&n32, sizeof(uint32_t), 1, out.f) == 1;
        std::vector<int32_t> iv((size_t)n);
        const int64_t C = P.columns, CR = (int64_t)P.columns * P.rows;
        for (int axis = 0; axis < 3 && ok; axis++) {
            for (int64_t h = 0; h < n; h++)
                iv[h] = (int32_t)(axis == 0 ? first[h] % C : axis == 1 ? (first[h] / C) % P.rows : first[h] / CR);
            ok = std::fwrite(iv.data(), sizeof(int32_t), iv.size(), out.f) == iv.size();
        }
        for (int k = 0; k < 10 && ok; k++) ok = std::fwrite(hf[k].data(), sizeof(float), hf[k].size(), out.f) == hf[k].size();
        if (ok) ok = std::fwrite(bin.data(), sizeof(int32_t), bin.size(), out.f) == bin.size();
        for (int64_t h = 0; h < n; h++) iv[h] = (int32_t)gantry[h];
        if (ok) ok = std::fwrite(iv.data(), sizeof(int32_t), iv.size(), out.f) == iv.size();
        if (!ok) return io_fail(ctx, "short write to " + path);
    }

    // hull and x_0 (export_hull :6540, export_x_0 :6551): raw bytes / floats
    const int64_t NV = P.num_voxels;
    if (ctx->hull) {
        std::vector<uint8_t> hull;
        PCT_TRY(copy_to_host(ctx, ctx->hull, NV, &hull));
        File out;
        out.f = std::fopen(join(dir, "hull.txt").c_str(), "wb");
        if (!out.f || std::fwrite(hull.data(), 1, hull.size(), out.f) != hull.size())
            return io_fail(ctx, "cannot write hull.txt in " + std::string(dir));
    }
    {
        std::vector<float> x0;
        PCT_TRY(copy_to_host(ctx, ctx->x0, NV, &x0));
        File out;
        out.f = std::fopen(join(dir, "x_0.txt").c_str(), "wb");
        if (!out.f || std::fwrite(x0.data(), sizeof(float), x0.size(), out.f) != x0.size())
            return io_fail(ctx, "cannot write x_0.txt in " + std::string(dir));
    }
    return PCT_OK;
}

extern "C" int pct_import_preprocessed(pct_ctx* ctx, const char* dir)
{
    if (!ctx || !dir) return PCT_E_ARG;
    const pct_params& P = ctx->P;
    const int64_t NV = P.num_voxels;
    hipStream_t s = ctx->stream;
    PCT_CHECK_HIP(ctx, hipStreamSynchronize(s));

    auto open = [&](const char* name, File* f) {
        f->f = std::fopen(join(dir, name).c_str(), "rb");
        return f->f != nullptr;
    };
    // voxels_per_path.bin fixes n and the path lengths
    File fc;
    if (!open("voxels_per_path.bin", &fc)) return io_fail(ctx, "voxels_per_path.bin missing in " + std::string(dir));
    const long long size_c = file_size(fc.f);
    uint32_t n32 = 0;
    if (std::fread(&n32, 4, 1, fc.f) != 1 || size_c != 4 + 4ll * n32)
        return io_fail(ctx, "voxels_per_path.bin: size does not match its history count");
    const int64_t n = n32;
    std::vector<uint32_t> cnt((size_t)n);
    if (n && std::fread(cnt.data(), 4, (size_t)n, fc.f) != (size_t)n) return io_fail(ctx, "voxels_per_path.bin: short read");
    std::vector<int64_t> off((size_t)n + 1, 0);
    for (int64_t h = 0; h < n; h++) {
        if (cnt[h] > max_path(P)) {
            ctx->err = "voxels_per_path.bin: a path of " + std::to_string(cnt[h]) + " voxels exceeds the limit " +
                       std::to_string(max_path(P)) + " (4 (COLUMNS + ROWS + SLICES) + 16)";
            return PCT_E_LIMIT;
        }
        off[h + 1] = off[h] + cnt[h];
    }
    const int64_t total = off[n];

    // per-history floats: chords (FP64 file preferred), WEPL
    std::vector<double> chord((size_t)n);
    {
        File f64;
        if (open("avg_chord_lengths_f64.bin", &f64)) {
            uint32_t m = 0;
            if (file_size(f64.f) != 4 + 8ll * n || std::fread(&m, 4, 1, f64.f) != 1 || m != n32 ||
                (n && std::fread(chord.data(), 8, (size_t)n, f64.f) != (size_t)n))
                return io_fail(ctx, "avg_chord_lengths_f64.bin does not hold the histories of voxels_per_path.bin");
        } else {
            File f;
            std::vector<float> c32((size_t)n);
            uint32_t m = 0;
            if (!open("avg_chord_lengths.bin", &f)) return io_fail(ctx, "avg_chord_lengths.bin missing in " + std::string(dir));
            if (file_size(f.f) != 4 + 4ll * n || std::fread(&m, 4, 1, f.f) != 1 || m != n32 ||
                (n && std::fread(c32.data(), 4, (size_t)n, f.f) != (size_t)n))
                return io_fail(ctx, "avg_chord_lengths.bin does not hold the histories of voxels_per_path.bin");
            for (int64_t h = 0; h < n; h++) chord[h] = (double)c32[h];
        }
    }
    std::vector<float> wepl((size_t)n);
    {
        File f;
        if (!open("WEPL.bin", &f)) return io_fail(ctx, "WEPL.bin missing in " + std::string(dir));
        const long long hdr = read_header(f.f, n32, 4ll * n, file_size(f.f));
        if (hdr < 0) return io_fail(ctx, "WEPL.bin does not hold the histories of voxels_per_path.bin");
        std::fseek(f.f, (long)hdr, SEEK_SET);
        if (n && std::fread(wepl.data(), 4, (size_t)n, f.f) != (size_t)n) return io_fail(ctx, "WEPL.bin: short read");
    }
    std::vector<float> x0((size_t)NV);
    {
        File f;
        if (!open("x_0.txt", &f)) return io_fail(ctx, "x_0.txt missing in " + std::string(dir));
        if (file_size(f.f) != 4ll * NV || std::fread(x0.data(), 4, (size_t)NV, f.f) != (size_t)NV)
            return io_fail(ctx, "x_0.txt does not hold NUM_VOXELS floats");
    }
    std::vector<uint8_t> hull;
    {
        File f;
        if (open("hull.txt", &f)) {
            hull.resize((size_t)NV);
            if (file_size(f.f) != NV || std::fread(hull.data(), 1, (size_t)NV, f.f) != (size_t)NV)
                return io_fail(ctx, "hull.txt does not hold NUM_VOXELS bytes");
            for (uint8_t& b : hull) b = b != 0;
        }
    }
    // histories.bin (optional): entry voxels, coordinates and angles, bins, gantry angles
    std::vector<int32_t> hv[3], bin((size_t)n, 0), gant((size_t)n, 0);
    std::vector<float> hf[10];
    bool have_hist = false;
    {
        File f;
        if (open("histories.bin", &f)) {
            uint32_t m = 0;
            if (file_size(f.f) != 4 + 15ll * 4 * n || std::fread(&m, 4, 1, f.f) != 1 || m != n32)
                return io_fail(ctx, "histories.bin does not hold the histories of voxels_per_path.bin");
            bool ok = true;
            for (int a = 0; a < 3 && ok; a++) {
                hv[a].resize((size_t)n);
                ok = !n || std::fread(hv[a].data(), 4, (size_t)n, f.f) == (size_t)n;
            }
            for (int k = 0; k < 10 && ok; k++) {
                hf[k].resize((size_t)n);
                ok = !n || std::fread(hf[k].data(), 4, (size_t)n, f.f) == (size_t)n;
            }
            if (ok) ok = !n || std::fread(bin.data(), 4, (size_t)n, f.f) == (size_t)n;
            if (ok) ok = !n || std::fread(gant.data(), 4, (size_t)n, f.f) == (size_t)n;
            if (!ok) return io_fail(ctx, "histories.bin: short read");
            for (int64_t h = 0; h < n; h++)
                if (hv[0][h] < 0 || hv[0][h] >= P.columns || hv[1][h] < 0 || hv[1][h] >= P.rows || hv[2][h] < 0 ||
                    hv[2][h] >= P.slices)
                    return io_fail(ctx, "histories.bin: an entry voxel lies outside the image");
            have_hist = true;
        }
    }

    // device state: per-history arrays, lists, images
    PCT_TRY(pct_ensure_hist_capacity(ctx, n));
    if (n > ctx->list_hist_cap || !ctx->list_off) {
        PCT_TRY(pct_alloc(ctx, &ctx->list_off, n + 1));
        PCT_TRY(pct_alloc(ctx, &ctx->list_cnt, n));
        PCT_TRY(pct_alloc(ctx, &ctx->chord, n));
        PCT_TRY(pct_alloc(ctx, &ctx->cap_scratch, n + 1));
        ctx->list_hist_cap = n;
    }
    if (total > ctx->list_cap) {
        PCT_TRY(pct_alloc(ctx, &ctx->list_vox, total));
        ctx->list_cap = total;
    }
    PCT_CHECK_HIP(ctx, hipMemcpy(ctx->list_off, off.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    if (n) {
        PCT_CHECK_HIP(ctx, hipMemcpy(ctx->list_cnt, cnt.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
        PCT_CHECK_HIP(ctx, hipMemcpy(ctx->chord, chord.data(), n * sizeof(double), hipMemcpyHostToDevice));
        PCT_CHECK_HIP(ctx, hipMemcpy(ctx->wepl, wepl.data(), n * sizeof(float), hipMemcpyHostToDevice));
        PCT_CHECK_HIP(ctx, hipMemcpy(ctx->bin_num, bin.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
        ctx->bin_perm_n = -1;
        std::vector<float> g(gant.begin(), gant.end());
        PCT_CHECK_HIP(ctx, hipMemcpy(ctx->gantry, g.data(), n * sizeof(float), hipMemcpyHostToDevice));
        if (!have_hist) {  // no entry voxels or coordinates: zeros, so that an export writes defined values
            PCT_CHECK_HIP(ctx, hipMemset(ctx->first_voxel, 0, n * sizeof(uint32_t)));
            for (int k = 0; k < 10; k++) PCT_CHECK_HIP(ctx, hipMemset(ctx->h[k], 0, n * sizeof(float)));
        } else {
            std::vector<uint32_t> first((size_t)n);
            for (int64_t h = 0; h < n; h++)
                first[h] = (uint32_t)(hv[0][h] + (int64_t)hv[1][h] * P.columns + (int64_t)hv[2][h] * P.columns * P.rows);
            PCT_CHECK_HIP(ctx, hipMemcpy(ctx->first_voxel, first.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
            for (int k = 0; k < 10; k++)
                PCT_CHECK_HIP(ctx, hipMemcpy(ctx->h[k], hf[k].data(), n * sizeof(float), hipMemcpyHostToDevice));
        }
    }
    // MLP.bin, streamed; every voxel index is checked before it reaches the device
    {
        File f;
        if (!open("MLP.bin", &f)) return io_fail(ctx, "MLP.bin missing in " + std::string(dir));
        const long long hdr = read_header(f.f, n32, 4ll * total, file_size(f.f));
        if (hdr < 0) return io_fail(ctx, "MLP.bin does not hold the paths of voxels_per_path.bin");
        std::fseek(f.f, (long)hdr, SEEK_SET);
        std::vector<uint32_t> buf;
        for (int64_t e = 0; e < total; e += kChunk) {
            const int64_t m = std::min(kChunk, total - e);
            buf.resize((size_t)m);
            if (std::fread(buf.data(), 4, (size_t)m, f.f) != (size_t)m) return io_fail(ctx, "MLP.bin: short read");
            for (int64_t q = 0; q < m; q++)
                if (buf[q] >= (uint64_t)NV) return io_fail(ctx, "MLP.bin: a voxel index is not below NUM_VOXELS");
            PCT_CHECK_HIP(ctx, hipMemcpy(ctx->list_vox + e, buf.data(), m * sizeof(uint32_t), hipMemcpyHostToDevice));
        }
    }
    PCT_TRY(pct_alloc(ctx, &ctx->x0, NV));
    PCT_TRY(pct_alloc(ctx, &ctx->x, NV));
    PCT_CHECK_HIP(ctx, hipMemcpy(ctx->x0, x0.data(), NV * sizeof(float), hipMemcpyHostToDevice));
    PCT_CHECK_HIP(ctx, hipMemcpy(ctx->x, x0.data(), NV * sizeof(float), hipMemcpyHostToDevice));
    if (!hull.empty()) {
        PCT_TRY(pct_alloc(ctx, &ctx->hull, NV));
        PCT_CHECK_HIP(ctx, hipMemcpy(ctx->hull, hull.data(), NV, hipMemcpyHostToDevice));
    } else {
        pct_free(&ctx->hull);
    }

    ctx->n_hist = ctx->n_recon = n;
    ctx->n_loaded = ctx->n_recon_volume = ctx->n_post_cuts = 0;
    ctx->list_entries = total;
    ctx->list_slots = total;
    ctx->drop_k = 0;
    ctx->tvs_L = 0;
    ctx->tvs_log.clear();
    ctx->stage = ST_LISTS;
    ctx->plan.vox_b_ready = false;
    ctx->plan.ready = false;
    PCT_TRY(pct_drop_prepare(ctx));
    return PCT_OK;
}
