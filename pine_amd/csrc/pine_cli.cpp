// `pine <scene.txt>`: the drop-in command line (src/cli/main.cpp:6-29) over the
// C ABI.  Loads the scene, renders it on the GPU and finalises the film the way
// pine's Film::Finalize does (film.cpp:45-100: divide by the filter weight,
// Uncharted2 filmic curve, gamma 1/2.2 with pine's pow, 8-bit BMP), and also
// writes the linear radiance as a PFM next to the BMP.
//
// Extra flags (after the scene file; pine takes exactly one argument):
//   --spp N      override samplesPerPixel
//   --no-write   skip the image files
#include "../../include/pinehip.h"
#include "pg_pstd.h"

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace pg;

namespace {

// Uncharted2Flimic (color.cpp:6-13)
f3 filmic(f3 v) {
    auto map = [](f3 x) {
        const f3 A = mk(0.15f, 0.15f, 0.15f), B = mk(0.50f, 0.50f, 0.50f), C = mk(0.10f, 0.10f, 0.10f),
                 D = mk(0.20f, 0.20f, 0.20f), E = mk(0.02f, 0.02f, 0.02f), F = mk(0.30f, 0.30f, 0.30f);
        return (x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F) - E / F;
    };
    return map(v * 2.0f) * mk(1, 1, 1) / map(mk(11.2f, 11.2f, 11.2f));
}

bool write_bmp(const std::string& path, int w, int h, const std::vector<float>& rgb) {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f)
        return false;
    const int row = (3 * w + 3) & ~3;
    const int filesize = 54 + row * h;
    unsigned char hdr[54] = {'B', 'M'};
    auto put32 = [&](int off, int v) {
        for (int k = 0; k < 4; k++)
            hdr[off + k] = (unsigned char)(v >> (8 * k));
    };
    put32(2, filesize);
    put32(10, 54);
    put32(14, 40);
    put32(18, w);
    put32(22, h);
    hdr[26] = 1;
    hdr[28] = 24;
    fwrite(hdr, 1, 54, f);
    std::vector<unsigned char> line(row, 0);
    // WriteImageBMP (fileio.cpp:88-160): bottom-up rows, BGR, value*256 clamped.
    for (int y = h - 1; y >= 0; y--) {
        for (int x = 0; x < w; x++)
            for (int c = 0; c < 3; c++) {
                float v = rgb[3 * ((size_t)y * w + x) + c] * 256.0f;
                v = v < 0.0f ? 0.0f : (v > 255.0f ? 255.0f : v);
                line[3 * x + (2 - c)] = (unsigned char)v;
            }
        fwrite(line.data(), 1, row, f);
    }
    fclose(f);
    return true;
}

bool write_pfm(const std::string& path, int w, int h, const std::vector<float>& rgb) {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f)
        return false;
    fprintf(f, "PF\n%d %d\n-1.0\n", w, h);
    for (int y = h - 1; y >= 0; y--)  // PFM stores the bottom row first
        fwrite(&rgb[3 * (size_t)y * w], sizeof(float), 3 * (size_t)w, f);
    fclose(f);
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        printf("Usage: pine [filename]\n");
        return 0;
    }
    int spp = 0;
    bool write = true;
    for (int i = 2; i < argc; i++) {
        if (!strcmp(argv[i], "--spp") && i + 1 < argc)
            spp = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--no-write"))
            write = false;
    }
    printf("[FileIO]Loading \"%s\"\n", argv[1]);
    pg_scene* scene = nullptr;
    if (pg_scene_load(argv[1], nullptr, &scene) != PG_OK) {
        fprintf(stderr, "[pine]%s\n", pg_last_error());
        return 1;
    }
    pg_render_options opts = {spp, 0};
    if (pg_render(scene, &opts, nullptr) != PG_OK) {
        fprintf(stderr, "[pine]%s\n", pg_last_error());
        pg_scene_release(scene);
        return 1;
    }
    pg_scene_info info;
    pg_scene_get_info(scene, &info);
    const int w = info.width, h = info.height, n = spp > 0 ? spp : info.spp;
    const double samples = (double)w * h * n;
    printf("[Rendering]Average:%.4fM Samples/s  (%dx%d, %d spp, %.3f s, %.1f Mrays/s)\n",
           samples / (info.last_render_ms * 1e3), w, h, n, info.last_render_ms * 1e-3,
           info.last_rays / (info.last_render_ms * 1e3));
    std::vector<float> film((size_t)w * h * 4);
    pg_scene_film(scene, film.data());
    char namebuf[1024];
    pg_scene_output_name(scene, namebuf, sizeof(namebuf));
    const std::string name = namebuf;
    pg_scene_release(scene);
    if (!write)
        return 0;
    std::vector<float> lin((size_t)w * h * 3), out((size_t)w * h * 3);
    for (size_t i = 0; i < (size_t)w * h; i++) {
        const float wt = film[4 * i + 3];
        f3 c = wt != 0.0f ? mk(film[4 * i] / wt, film[4 * i + 1] / wt, film[4 * i + 2] / wt)
                          : mk(0, 0, 0);
        lin[3 * i] = c.x, lin[3 * i + 1] = c.y, lin[3 * i + 2] = c.z;
        f3 t = filmic(c);
        out[3 * i + 0] = pstd::pow(t.x, 1.0f / 2.2f);
        out[3 * i + 1] = pstd::pow(t.y, 1.0f / 2.2f);
        out[3 * i + 2] = pstd::pow(t.z, 1.0f / 2.2f);
    }
    // Film outputFileName (film.cpp:9); pine only writes BMP, so the extension
    // becomes .bmp.
    size_t dot = name.find_last_of('.');
    std::string stem = dot == std::string::npos ? name : name.substr(0, dot);
    if (!write_bmp(stem + ".bmp", w, h, out) || !write_pfm(stem + ".pfm", w, h, lin)) {
        fprintf(stderr, "[pine]cannot write %s.bmp\n", stem.c_str());
        return 1;
    }
    printf("[Film]wrote %s.bmp and %s.pfm\n", stem.c_str(), stem.c_str());
    return 0;
}
