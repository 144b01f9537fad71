// Exercises the point-cloud and image drop-in headers the way the reference's
// tests/Object_test.cpp does: objread<float> -> objset (two ranges, ZMIDDLE) ->
// objinfo -> objsort -> phase-only Cgh -> phase8bit -> bmpwrite, plus the double
// cloud, the offset/scale reader, quant8bit, complex2img / complex2real and Save.
// Every intermediate is written to <dir> for the Python test to check against the
// oracle.
//   ingest_test <dir> <file.3d>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "openol.h"

template <typename T>
static void save(const std::string& path, const T* p, size_t count) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(p), sizeof(T) * count);
}

template <typename T>
static void save_obj(const std::string& path, ol::Object<T>& obj) {
    std::vector<T> v((size_t)obj.size * 3);
    for (int j = 0; j < obj.size; j++) {
        v[3 * j] = obj[j].x;
        v[3 * j + 1] = obj[j].y;
        v[3 * j + 2] = obj[j].z;
    }
    save(path, v.data(), v.size());
}

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const std::string dir = argv[1];
    const char* path3d = argv[2];
    typedef float PREC_T;  // the reference's test precision
    const int WIDTH = 256, HEIGHT = 128;
    PREC_T p = 3.74e-6, lambda = 532e-9;
    try {
        // tests/Object_test.cpp:18-28
        ol::Vect3<PREC_T> offset((float)(WIDTH / 2 * p), (float)(HEIGHT / 2 * p), 1.0f);
        auto obj = ol::objread<PREC_T>(path3d);
        save_obj(dir + "/obj_read.bin", obj);
        ol::objset(obj, WIDTH * p * 0.8, HEIGHT * p * 0.8, offset, ol::ZMIDDLE);
        save_obj(dir + "/obj_set.bin", obj);
        std::cout << obj.size << std::endl;
        ol::objinfo(obj);
        ol::objsort(obj);
        save_obj(dir + "/obj_sort.bin", obj);
        std::cout << "z0 " << obj[0].z << std::endl;

        // double cloud, single-range objset with ZMIN (prints its bounds)
        auto objd = ol::objread<double>(path3d);
        ol::objset(objd, 1e-3, ol::Vect3<double>(1e-4, -2e-4, 0.1), ol::ZMIN);
        save_obj(dir + "/objd_set.bin", objd);
        // device-resident cloud: sort in HBM
        auto g = ol::Object2dev(objd);
        ol::objsort(g);
        ol::Object<double> back(objd.size);
        {
            std::vector<olx_point> v((size_t)objd.size);
            ol::detail::check(olx_copy_d2h(ol::detail::ctx(), v.data(), g.points.get(),
                                           sizeof(olx_point) * v.size()), "copy");
            for (int j = 0; j < objd.size; j++) back[j] = ol::Vect3<double>(v[j].x, v[j].y, v[j].z);
        }
        save_obj(dir + "/objd_gsort.bin", back);
        // reader with offset and scale (float offsets, double scale)
        float offs[3] = {0.5f, -1.25f, 3.0f};
        auto objo = ol::objread<double>(path3d, offs, 0.001);
        save_obj(dir + "/obj_offscale.bin", objo);

        // tests/Object_test.cpp:32-42
        std::unique_ptr<PREC_T[]> phase = std::make_unique<PREC_T[]>(WIDTH * HEIGHT);
        auto img = std::make_unique<uint8_t[]>(WIDTH * HEIGHT);
        ol::Cgh(obj, phase, HEIGHT, WIDTH, p, p, lambda);
        save(dir + "/phase.bin", phase.get(), (size_t)WIDTH * HEIGHT);
        ol::phase8bit(phase, img, HEIGHT, WIDTH);
        save(dir + "/phase8.bin", img.get(), (size_t)WIDTH * HEIGHT);
        ol::bmpwrite((dir + "/olObject.bmp").c_str(), img, HEIGHT, WIDTH);
        auto rd = std::make_unique<uint8_t[]>(WIDTH * HEIGHT);
        ol::bmpread((dir + "/olObject.bmp").c_str(), rd, HEIGHT, WIDTH);
        if (std::memcmp(rd.get(), img.get(), (size_t)WIDTH * HEIGHT) != 0) std::cout << "bmp mismatch\n";
        ol::quant8bit(phase, img, HEIGHT, WIDTH);
        save(dir + "/quant8.bin", img.get(), (size_t)WIDTH * HEIGHT);
        ol::Save((dir + "/quant.bmp").c_str(), phase, HEIGHT, WIDTH);

        // complex hologram and its encodings
        auto u = std::make_unique<std::complex<PREC_T>[]>(WIDTH * HEIGHT);
        ol::Cgh(obj, u, HEIGHT, WIDTH, p, lambda);
        save(dir + "/u.bin", u.get(), (size_t)WIDTH * HEIGHT);
        ol::Save((dir + "/amp.bmp").c_str(), u, HEIGHT, WIDTH, ol::AMP);
        ol::complex2img(u, img, HEIGHT, WIDTH, true);
        save(dir + "/img_legacy_real.bin", img.get(), (size_t)WIDTH * HEIGHT);
        ol::complex2img(u, img, HEIGHT, WIDTH, ol::PHASE);
        save(dir + "/img_phase.bin", img.get(), (size_t)WIDTH * HEIGHT);
        auto re = ol::complex2real(u, HEIGHT, WIDTH, ol::AMP);
        save(dir + "/real_amp.bin", re.get(), (size_t)WIDTH * HEIGHT);
        auto du = ol::cpu2dev(u, (size_t)WIDTH * HEIGHT);
        ol::Save((dir + "/dev_imag.bmp").c_str(), du, HEIGHT, WIDTH, ol::IMAG);
        try {
            ol::Save((dir + "/x.png").c_str(), u, HEIGHT, WIDTH);
        } catch (const std::exception& e) {
            std::cout << "caught: " << e.what() << std::endl;
        }
        ol::synchronize();
    } catch (const std::exception& e) {
        std::cout << "error: " << e.what() << std::endl;
        return 1;
    }
    std::cout << "done" << std::endl;
    return 0;
}
