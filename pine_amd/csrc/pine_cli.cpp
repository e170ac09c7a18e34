// `pine <scene.txt>`: the drop-in command line (src/cli/main.cpp:6-29) over the
// C ABI.  Loads the scene, renders it on the GPU and finalises the film the way
// pine's Film::Finalize does (film.cpp:45-100): divide by the filter weight,
// the Uncharted2 filmic curve unless the film says applyToneMapping: false,
// gamma 1/2.2 with pine's pow, and an 8-bit BMP when outputFileName ends in
// .bmp (pine's SaveImage writes nothing for any other extension and warns,
// fileio.cpp:174-180).  The linear radiance is also written as a PFM next to
// the output name (an addition: pine has no HDR output).
//
// Extra flags (after the scene file; pine takes exactly one argument):
//   --spp N        override samplesPerPixel
//   --no-write     skip the image files
//   --film PATH    also write the film's accumulators (i32 w, i32 h, then
//                  h*w*4 floats: sum w*r, sum w*g, sum w*b, sum w; row 0 on top)
// and a mode that needs no scene and no GPU:
//   pine --finalize FILM OUT.bmp [--no-tone-mapping]
//                  Film::Finalize of a film file written by --film.
#include "../../include/pinehip.h"
#include "pg_pstd.h"

#include <cstdio>
#include <cstdlib>
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

// SaveImage + WriteImageBMP (fileio.cpp:88-184): each channel
// clamp(v * 256, 0, 255) (pstd::clamp = min(max(v, 0), 255), so NaN gives 0)
// truncated to 8 bits; BGR; bottom row first; header fields as pine writes
// them (bfSize = 54 + 3*w*h).  Rows of widths that are not a multiple of 4 are
// padded to 4 bytes here; pine's writer reads those rows from the wrong offset.
bool write_bmp(const std::string& path, int w, int h, const std::vector<float>& rgb) {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f)
        return false;
    const int pad = (4 - (3 * w) % 4) % 4;
    const int filesize = 54 + 3 * w * h;
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
    std::vector<unsigned char> line(3 * w + pad, 0);
    for (int y = h - 1; y >= 0; y--) {
        for (int x = 0; x < w; x++)
            for (int c = 0; c < 3; c++) {
                float v = rgb[3 * ((size_t)y * w + x) + c] * 256.0f;
                v = v > 0.0f ? v : 0.0f;
                v = v < 255.0f ? v : 255.0f;
                line[3 * x + (2 - c)] = (unsigned char)v;
            }
        fwrite(line.data(), 1, line.size(), f);
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

// Film::CopyToRGBArray, ApplyToneMapping and ApplyGammaCorrection
// (film.cpp:72-100) of a film of h*w*4 accumulators; lin gets rgb / weight.
void finalize(const std::vector<float>& film, size_t n, bool tone_map, std::vector<float>& lin,
              std::vector<float>& out) {
    lin.assign(3 * n, 0.0f);
    out.assign(3 * n, 0.0f);
    for (size_t i = 0; i < n; i++) {
        const float wt = film[4 * i + 3];
        f3 c = wt != 0.0f ? mk(film[4 * i] / wt, film[4 * i + 1] / wt, film[4 * i + 2] / wt)
                          : mk(0, 0, 0);
        lin[3 * i] = c.x, lin[3 * i + 1] = c.y, lin[3 * i + 2] = c.z;
        if (tone_map)
            c = filmic(c);
        out[3 * i + 0] = pstd::pow(c.x, 1.0f / 2.2f);
        out[3 * i + 1] = pstd::pow(c.y, 1.0f / 2.2f);
        out[3 * i + 2] = pstd::pow(c.z, 1.0f / 2.2f);
    }
}

bool read_film(const char* path, int& w, int& h, std::vector<float>& film) {
    FILE* f = fopen(path, "rb");
    if (!f)
        return false;
    int32_t hdr[2];
    bool ok = fread(hdr, 4, 2, f) == 2 && hdr[0] > 0 && hdr[1] > 0;
    if (ok) {
        w = hdr[0], h = hdr[1];
        film.resize((size_t)w * h * 4);
        ok = fread(film.data(), sizeof(float), film.size(), f) == film.size();
    }
    fclose(f);
    return ok;
}

bool write_film(const std::string& path, int w, int h, const std::vector<float>& film) {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f)
        return false;
    int32_t hdr[2] = {w, h};
    bool ok = fwrite(hdr, 4, 2, f) == 2 && fwrite(film.data(), sizeof(float), film.size(), f) == film.size();
    fclose(f);
    return ok;
}

bool ends_with(const std::string& s, const char* suffix) {
    const size_t n = strlen(suffix);
    return s.size() >= n && s.compare(s.size() - n, n, suffix) == 0;
}

int finalize_mode(int argc, char** argv) {
    if (argc < 4) {
        printf("Usage: pine --finalize [film] [output.bmp] [--no-tone-mapping]\n");
        return 2;
    }
    const bool tone_map = !(argc > 4 && !strcmp(argv[4], "--no-tone-mapping"));
    int w = 0, h = 0;
    std::vector<float> film, lin, out;
    if (!read_film(argv[2], w, h, film)) {
        fprintf(stderr, "[pine]cannot read film %s\n", argv[2]);
        return 1;
    }
    finalize(film, (size_t)w * h, tone_map, lin, out);
    if (!write_bmp(argv[3], w, h, out)) {
        fprintf(stderr, "[pine]cannot write %s\n", argv[3]);
        return 1;
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        printf("Usage: pine [filename]\n");
        return 0;
    }
    if (!strcmp(argv[1], "--finalize"))
        return finalize_mode(argc, argv);
    int spp = 0;
    bool write = true;
    std::string film_path;
    for (int i = 2; i < argc; i++) {
        if (!strcmp(argv[i], "--spp") && i + 1 < argc)
            spp = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--no-write"))
            write = false;
        else if (!strcmp(argv[i], "--film") && i + 1 < argc)
            film_path = argv[++i];
    }
    printf("[FileIO]Loading \"%s\"\n", argv[1]);
    pg_scene* scene = nullptr;
    if (pg_scene_load(argv[1], nullptr, &scene) != PG_OK) {
        fprintf(stderr, "[pine]%s\n", pg_last_error());
        return 1;
    }
    pg_render_options opts = {spp, 0, 0, 0, 0};
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
    if (!film_path.empty() && !write_film(film_path, w, h, film)) {
        fprintf(stderr, "[pine]cannot write %s\n", film_path.c_str());
        return 1;
    }
    if (!write)
        return 0;
    std::vector<float> lin, out;
    finalize(film, (size_t)w * h, info.apply_tone_mapping != 0, lin, out);
    // Film outputFileName (film.cpp:9): pine's SaveImage writes BMP files only.
    const size_t dot = name.find_last_of('.');
    const std::string stem = dot == std::string::npos ? name : name.substr(0, dot);
    if (ends_with(name, ".bmp")) {
        if (!write_bmp(name, w, h, out)) {
            fprintf(stderr, "[pine]cannot write %s\n", name.c_str());
            return 1;
        }
        printf("[Film]wrote %s\n", name.c_str());
    } else {
        printf("[Warning]%s has unsupported image file extension\n", name.c_str());
    }
    if (!write_pfm(stem + ".pfm", w, h, lin)) {
        fprintf(stderr, "[pine]cannot write %s.pfm\n", stem.c_str());
        return 1;
    }
    printf("[Film]wrote %s.pfm (linear)\n", stem.c_str());
    return 0;
}
