// Host code of libolx.so under AddressSanitizer and UndefinedBehaviorSanitizer, on
// the CPU (no device): the point-cloud and BMP file readers on the reference's
// fixtures and on truncated / malformed files, the BMP writer, the host-side checks
// and plane / layer rules, and the argument validation of the entry points (each
// must fail with a message before touching memory or the device).  Built from the
// library's sources with the sanitizers on their host side (device code as usual)
// by tests/cpp/Makefile.san.
//   host_san_test <golden dir> <scratch dir>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "olx.h"

static int failures = 0;
#define EXPECT(cond, what)                                                            \
    do {                                                                              \
        if (!(cond)) {                                                                \
            std::cerr << "FAILED: " << what << " (" << __FILE__ << ":" << __LINE__ << ")\n"; \
            failures++;                                                               \
        }                                                                             \
    } while (0)

static void write_bytes(const std::string& path, const std::vector<uint8_t>& b) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(b.data()), (std::streamsize)b.size());
}

static std::vector<uint8_t> read_bytes(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    return std::vector<uint8_t>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

// a call rejected with a negative code and a non-empty message
static void rejected(int rc, const char* what) {
    EXPECT(rc < 0, what);
    const char* m = olx_last_error();
    EXPECT(m && std::strlen(m) > 0, what);
}

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const std::string gold = argv[1], tmp = argv[2];
    EXPECT(olx_version() > 0, "version");
    // ---- point clouds: the reference's fixtures, counted then read
    for (const char* name : {"cube284.3d", "horn02-f000.3d"}) {
        const std::string path = gold + "/" + name;
        int64_t n = -1;
        EXPECT(olx_obj_read(path.c_str(), OLX_OBJ_AUTO, nullptr, 0, &n, 1.0, 0, 0, 0) == 0, name);
        EXPECT(n > 0, name);
        std::vector<olx_point> v((size_t)n);
        int64_t m = -1;
        EXPECT(olx_obj_read(path.c_str(), OLX_OBJ_3D, v.data(), n, &m, 2e-6, 1e-3, 0, 0.1) == 0, name);
        EXPECT(m == n, name);
        // a capacity below N writes only cap points
        std::vector<olx_point> w(3);
        EXPECT(olx_obj_read(path.c_str(), OLX_OBJ_3D, w.data(), 3, &m, 1.0, 0, 0, 0) == 0, name);
        // the same file truncated mid-point, with a count larger than its points,
        // and with a negative count: errors
        auto bytes = read_bytes(path);
        for (size_t cut : {size_t(2), size_t(4), size_t(4 + 12 * 2 + 5), bytes.size() - 1}) {
            const std::string t = tmp + "/trunc.3d";
            write_bytes(t, std::vector<uint8_t>(bytes.begin(), bytes.begin() + (long)std::min(cut, bytes.size())));
            int64_t k = 0;
            std::vector<olx_point> big((size_t)n);
            rejected(olx_obj_read(t.c_str(), OLX_OBJ_3D, big.data(), n, &k, 1.0, 0, 0, 0), "truncated .3d");
        }
        std::vector<uint8_t> neg = bytes;
        const int32_t minus = -5;
        std::memcpy(neg.data(), &minus, 4);
        write_bytes(tmp + "/neg.3d", neg);
        int64_t k = 0;
        rejected(olx_obj_read((tmp + "/neg.3d").c_str(), OLX_OBJ_3D, nullptr, 0, &k, 1, 0, 0, 0),
                 "negative count");
    }
    // a float cloud written here
    {
        std::vector<uint8_t> b(4 + 12 * 5);
        const int32_t n = 5;
        std::memcpy(b.data(), &n, 4);
        for (int i = 0; i < 15; i++) {
            const float f = 0.25f * (float)i - 1.0f;
            std::memcpy(b.data() + 4 + 4 * i, &f, 4);
        }
        write_bytes(tmp + "/c.3df", b);
        std::vector<olx_point> v(5);
        int64_t m = 0;
        EXPECT(olx_obj_read((tmp + "/c.3df").c_str(), OLX_OBJ_AUTO, v.data(), 5, &m, 1, 0, 0, 0) == 0, ".3df");
        EXPECT(m == 5 && std::fabs(v[4].z - 2.5) < 1e-12, ".3df values");
    }
    int64_t k = 0;
    rejected(olx_obj_read((tmp + "/missing.3d").c_str(), OLX_OBJ_3D, nullptr, 0, &k, 1, 0, 0, 0),
             "missing file");
    rejected(olx_obj_read(nullptr, OLX_OBJ_3D, nullptr, 0, &k, 1, 0, 0, 0), "null path");
    // ---- BMP: round trips at widths with every row padding, then a short file
    for (int nx : {1, 2, 3, 4, 5, 17, 64}) {
        const int ny = 7;
        std::vector<uint8_t> img((size_t)ny * nx), back((size_t)ny * nx);
        for (size_t i = 0; i < img.size(); i++) img[i] = (uint8_t)(i * 37 + 11);
        const std::string path = tmp + "/img.bmp";
        EXPECT(olx_bmp_write(path.c_str(), img.data(), ny, nx) == 0, "bmp write");
        EXPECT(olx_bmp_read(path.c_str(), back.data(), ny, nx) == 0, "bmp read");
        EXPECT(img == back, "bmp round trip");
        // short by the last row's padding (nx % 4 bytes) and one data byte
        auto bytes = read_bytes(path);
        const long cut = nx % 4 + 1;
        write_bytes(tmp + "/short.bmp", std::vector<uint8_t>(bytes.begin(), bytes.end() - cut));
        rejected(olx_bmp_read((tmp + "/short.bmp").c_str(), back.data(), ny, nx), "short bmp");
    }
    std::vector<uint8_t> one(4);
    rejected(olx_bmp_write((tmp + "/x.bmp").c_str(), one.data(), 0, 4), "bmp size");
    // ---- host-side rules
    EXPECT(olx_asm_prop_check(1024, 1024, 8e-6, 8e-6, 532e-9, 0.1) == 1, "asm check");
    EXPECT(olx_fresnel_prop_check(1024, 1024, 8e-6, 8e-6, 532e-9, 0.1) == 0, "fresnel check");
    const double pz = olx_wrp_plane_z(4096, 4096, 8e-6, 532e-9, 0.151, 8e-6, -1.0f, 0);
    EXPECT(pz < 0.151, "plane rule");
    // (0.16 - 0.1) / 0.1 * 8 = 4.8: layer 4, centre 0.1 + 4.5 x 0.0125
    EXPECT(std::fabs(olx_layer_depth(0.16, 0.1, 0.2, 8) - 0.15625) < 1e-15, "layer depth");
    // ---- argument validation: every entry point rejects a null context first
    float f[8] = {0};
    olx_point pt{0, 0, 0.1, 1, 0};
    rejected(olx_asm_c64(nullptr, f, 4, 4, 8e-6, 8e-6, 532e-9, 0.1, 0), "asm null ctx");
    rejected(olx_asm_c64_noalloc(nullptr, f, 4, 4, 8e-6, 8e-6, 532e-9, 0.1, 0), "noalloc null ctx");
    rejected(olx_fresnel_c64(nullptr, f, 4, 4, 8e-6, 8e-6, 532e-9, 0.1, 0), "fresnel null ctx");
    rejected(olx_arss_c64(nullptr, f, f, 4, 4, 8e-6, 8e-6, 532e-9, 0.1, 0.5, 0, 0, 0), "arss");
    rejected(olx_prop_batch_c64(nullptr, OLX_MODE_ASM, f, 2, 4, 4, 8e-6, 8e-6, 532e-9, 0.1, 0,
                                nullptr, nullptr, nullptr, 0), "batch");
    rejected(olx_asm_c128(nullptr, f, 4, 4, 8e-6, 8e-6, 532e-9, 0.1, 0), "asm c128");
    rejected(olx_fft2d_c128(nullptr, f, 4, 4, -1), "fft c128");
    rejected(olx_cgh_c64(nullptr, &pt, 1, f, nullptr, nullptr, 2, 2, 8e-6, 8e-6, 532e-9, 0), "cgh");
    rejected(olx_cgh_img8_c64(nullptr, &pt, 1, nullptr, (uint8_t*)f, 2, 2, 8e-6, 8e-6, 532e-9, 0), "img8");
    rejected(olx_wrp_c64(nullptr, &pt, 1, f, 2, 2, 8e-6, 532e-9, 0.05, 8e-6, 0), "wrp");
    rejected(olx_cgh_layered_c64(nullptr, &pt, 1, f, 2, 2, 8e-6, 8e-6, 532e-9, 4, 0.05, 0.2, 0, 0),
             "layered");
    rejected(olx_window_copy(nullptr, f, 2, 2, f, 2, 2, 0, 0, 8), "window");
    rejected(olx_mul_c64(nullptr, f, f, f, 4), "mul");
    rejected(olx_phase8bit(nullptr, f, (uint8_t*)f, 4, 0), "phase8bit");
    rejected(olx_lens_tile_c64(nullptr, f, 4, f, 4, 4, 0, 0, 8e-6f, 0.01f, 532e-9f, 0.f, 0), "lens tile");
    if (failures) {
        std::cerr << failures << " failures\n";
        return 1;
    }
    std::cout << "host sanitize run: ok\n";
    return 0;
}
