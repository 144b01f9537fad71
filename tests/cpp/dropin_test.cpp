// Exercises the drop-in headers (include/openol.h) the way the reference's test
// mains call ol:: functions (tests/Prop_test.cpp, tests/gProp_test.cu,
// tests/gObject_test.cu), writing every result to <dir>/<name>.bin for the
// Python test to compare against the oracle.
//   dropin_test <dir>   reads <dir>/field.bin (complex64 n x n), <dir>/points.bin
//                       (m x 3 double), <dir>/params.txt ("n m")
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
static std::vector<T> load(const std::string& path, size_t count) {
    std::vector<T> v(count);
    std::ifstream f(path, std::ios::binary);
    f.read(reinterpret_cast<char*>(v.data()), sizeof(T) * count);
    if (!f) throw std::runtime_error("short read: " + path);
    return v;
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string dir = argv[1];
    int n = 0, m = 0;
    {
        std::ifstream f(dir + "/params.txt");
        f >> n >> m;
    }
    const double p = 8e-6, lambda = 532e-9;
    const auto field = load<std::complex<float>>(dir + "/field.bin", (size_t)n * n);
    const auto xyz = load<double>(dir + "/points.bin", (size_t)m * 3);
    auto fresh = [&] {
        auto u = std::make_unique<std::complex<float>[]>((size_t)n * n);
        for (size_t i = 0; i < (size_t)n * n; i++) u[i] = field[i];
        return u;
    };
    try {
        // host API, complex<float> (fftw3f path), PREC_T = double
        auto u = fresh();
        std::complex<float>* before = u.get();
        ol::AsmProp(u, n, n, p, p, lambda, 0.05);
        if (u.get() != before) std::cout << "allocation changed\n";
        save(dir + "/asm.bin", u.get(), (size_t)n * n);
        // complex<double> (fftw3 path)
        auto ud = std::make_unique<std::complex<double>[]>((size_t)n * n);
        for (size_t i = 0; i < (size_t)n * n; i++) ud[i] = std::complex<double>(field[i]);
        ol::AsmProp(ud, n, n, p, p, lambda, 0.05);
        save(dir + "/asm_c128.bin", ud.get(), (size_t)n * n);
        auto fresh_d = [&] {
            auto v = std::make_unique<std::complex<double>[]>((size_t)n * n);
            for (size_t i = 0; i < (size_t)n * n; i++) v[i] = std::complex<double>(field[i]);
            return v;
        };
        ud = fresh_d();
        ol::FresnelProp(ud, n, n, p, p, lambda, 0.2);
        save(dir + "/fresnel_c128.bin", ud.get(), (size_t)n * n);
        ud = fresh_d();
        ol::FresnelPropBandLimit(ud, n, n, p, p, lambda, 0.002);
        save(dir + "/fresnel_bl_c128.bin", ud.get(), (size_t)n * n);
        ud = fresh_d();
        ol::Prop(ud, n, n, p, p, lambda, 0.05, ol::AUTO);
        save(dir + "/prop_c128.bin", ud.get(), (size_t)n * n);
        ud = fresh_d();
        ol::mAsmProp(ud, n, n, p, p, lambda, 0.05);
        save(dir + "/masm_c128.bin", ud.get(), (size_t)n * n);
        u = fresh();
        ol::FresnelProp(u, n, n, p, p, lambda, 0.2);
        save(dir + "/fresnel.bin", u.get(), (size_t)n * n);
        u = fresh();
        ol::FresnelPropBandLimit(u, n, n, p, p, lambda, 0.002);
        save(dir + "/fresnel_bl.bin", u.get(), (size_t)n * n);
        u = fresh();
        ol::Prop(u, n, n, p, p, lambda, 0.05, ol::AUTO);
        save(dir + "/prop.bin", u.get(), (size_t)n * n);
        u = fresh();
        ol::shiftedFresnelProp(u, n, n, p, p, lambda, 0.1, 0.5, 1e-4, -2e-4);
        save(dir + "/shifted.bin", u.get(), (size_t)n * n);
        std::cout << "AsmPropCheck " << ol::AsmPropCheck(n, n, p, p, lambda, 0.05) << "\n";
        // off-axis and split propagation (PREC_T = float, as the reference's tests)
        {
            const float pf = (float)p, lf = (float)lambda;
            auto us = fresh();
            std::unique_ptr<std::complex<float>[]> ud;
            ol::shiftedAsmProp(us, ud, n, n, pf, pf, lf, 0.004f, 0.0f, 0.5f * n * pf);
            save(dir + "/shifted_asm.bin", ud.get(), (size_t)n * n);
            u = fresh();
            ol::splitAsmProp(u, n, n, pf, pf, lf, 0.002f);
            save(dir + "/split_cpu.bin", u.get(), (size_t)n * n);
            u = fresh();
            ol::gsplitProp(u, n, n, pf, pf, lf, 0.002f, ol::ASM);
            save(dir + "/gsplit.bin", u.get(), (size_t)n * n);
            u = fresh();
            ol::gsplitProp(u, n, n, pf, pf, lf, 0.05f, ol::FRESNEL, 2, 4);
            save(dir + "/gsplit_fresnel.bin", u.get(), (size_t)n * n);
            u = fresh();
            ol::mAsmProp(u, n, n, p, p, lambda, 0.05);
            save(dir + "/masm.bin", u.get(), (size_t)n * n);
            u = fresh();
            std::unique_ptr<std::complex<float>[]> F;
            ol::fft2d_hyblid(u, F, n, n);
            save(dir + "/hfft.bin", F.get(), (size_t)n * n);
            ol::ifft2d_hyblid(F, F, n, n);
            save(dir + "/hifft.bin", F.get(), (size_t)n * n);
            // tilted planes (the reference's tests/tiltedProp_test.cpp call)
            u = fresh();
            ol::TiltedReferenceProp<float>(u, n, n, pf, pf, lf, 0.004f, (float)(M_PI / 180 * 2),
                                           ol::BICUBIC, ol::ASM);
            save(dir + "/tilted_ref.bin", u.get(), (size_t)n * n);
            u = fresh();
            ol::TiltedSourceProp<float>(u, n, n, pf, pf, lf, 0.004f, (float)(M_PI / 180 * -3));
            save(dir + "/tilted_src.bin", u.get(), (size_t)n * n);
            // optical elements (include/Diffuser.h, include/Lensarry.h)
            u = fresh();
            ol::Diffuser<float> dif(n, n);
            dif.random_phase(u, ol::PLAY, 3, (float)(2 * M_PI), 1, -2);
            save(dir + "/diffuser.bin", u.get(), (size_t)n * n);
            u = fresh();
            ol::Lensarry<float> lens(n, n, pf, 16 * pf, 0.02f, lf);
            lens.mult_complex(u, ol::RECORD, 2, 3);
            save(dir + "/lens_mult.bin", u.get(), (size_t)n * n);
            ol::LensarryCircle<float> circ(n, n, pf, 16 * pf, 0.02f, lf);
            circ.gen_complex_limit(40e-6f);
            save(dir + "/lens_circle.bin", circ.u.get(), (size_t)n * n);
        }

        // CGH on a host point cloud (ol::Object<double>)
        ol::Object<double> obj(m);
        for (int j = 0; j < m; j++) obj[j] = ol::Vect3<double>(xyz[3 * j], xyz[3 * j + 1], xyz[3 * j + 2]);
        auto U = std::make_unique<std::complex<float>[]>((size_t)n * n);
        ol::Cgh(obj, U, n, n, p, lambda);
        save(dir + "/cgh.bin", U.get(), (size_t)n * n);
        auto ph = std::make_unique<float[]>((size_t)n * n);
        ol::Cgh(obj, ph, n, n, p, p, lambda);
        save(dir + "/cgh_phase.bin", ph.get(), (size_t)n * n);
        // normalized variant: the same cloud in pixel units
        ol::Object<double> pix(m);
        for (int j = 0; j < m; j++)
            pix[j] = ol::Vect3<double>(xyz[3 * j] / p, xyz[3 * j + 1] / p, xyz[3 * j + 2] / p);
        ol::Cgh(pix, ph, n, n, p, lambda, true);
        save(dir + "/cgh_norm.bin", ph.get(), (size_t)n * n);

        // device API
        auto d = ol::cpu2dev(fresh(), (size_t)n * n);
        ol::gAsmProp(d, n, n, p, p, lambda, 0.05);
        auto hd = ol::dev2cpu(d, (size_t)n * n);
        save(dir + "/gasm.bin", hd.get(), (size_t)n * n);
        auto src = ol::cpu2dev(fresh(), (size_t)n * n);
        ol::dfield dst;
        ol::ARSSFresnelProp(src, dst, n, n, p, p, lambda, 0.15, 0.5);
        hd = ol::dev2cpu(dst, (size_t)n * n);
        save(dir + "/arss.bin", hd.get(), (size_t)n * n);
        ol::dfield gs;
        ol::gshiftedAsmProp(src, gs, n, n, p, p, lambda, 0.004, -0.25 * n * p, 0.5 * n * p);
        hd = ol::dev2cpu(gs, (size_t)n * n);
        save(dir + "/gshifted.bin", hd.get(), (size_t)n * n);
        ol::shiftedProp(src, gs, n, n, p, p, lambda, 0.05, 1e-4, 0.0, ol::FRESNEL);
        hd = ol::dev2cpu(gs, (size_t)n * n);
        save(dir + "/shiftedprop_fresnel.bin", hd.get(), (size_t)n * n);
        auto gobj = ol::Object2dev(obj);
        auto dU = ol::make_device<std::complex<float>[]>((size_t)n * n);
        ol::gCgh(gobj, dU, n, n, p, lambda);
        hd = ol::dev2cpu(dU, (size_t)n * n);
        save(dir + "/gcgh.bin", hd.get(), (size_t)n * n);
        ol::WRPStep1(gobj, dU, n, n, p, lambda, 0.15);
        hd = ol::dev2cpu(dU, (size_t)n * n);
        save(dir + "/wrp.bin", hd.get(), (size_t)n * n);
        double pz = ol::WRPMethod(gobj, dU, n, n, p, lambda, 0.151, 0.99f);
        hd = ol::dev2cpu(dU, (size_t)n * n);
        save(dir + "/wrpm.bin", hd.get(), (size_t)n * n);
        std::printf("plane_z %.17g\n", pz);

        // tables
        auto H = std::make_unique<std::complex<float>[]>((size_t)2 * n * 2 * n);
        ol::AsmTransferF(H, 2 * n, 2 * n, 1 / (p * 2 * n), 1 / (p * 2 * n), lambda, 0.05);
        save(dir + "/H.bin", H.get(), (size_t)4 * n * n);
        auto h = std::make_unique<std::complex<float>[]>((size_t)2 * n * 2 * n);
        ol::FresnelResponseBandLimit(h, 2 * n, 2 * n, p, p, lambda, 0.002);
        save(dir + "/h.bin", h.get(), (size_t)4 * n * n);
        const double dq = 1 / (p * 2 * n);
        ol::shiftedAsmTransferF(H, 2 * n, 2 * n, dq, dq, lambda, 0.05, 2e-4, -5e-4);
        save(dir + "/Hs.bin", H.get(), (size_t)4 * n * n);
        auto gH = ol::make_device<std::complex<float>[]>((size_t)4 * n * n);
        ol::gAsmTransferF(gH, 2 * n, 2 * n, dq, dq, lambda, 0.05);
        save(dir + "/gH.bin", ol::dev2cpu(gH, (size_t)4 * n * n).get(), (size_t)4 * n * n);
        ol::gshiftedAsmTransferF(gH, 2 * n, 2 * n, dq, dq, lambda, 0.05, 2e-4, -5e-4);
        save(dir + "/gHs.bin", ol::dev2cpu(gH, (size_t)4 * n * n).get(), (size_t)4 * n * n);
        ol::gFresnelResponse(gH, 2 * n, 2 * n, p, p, lambda, 0.002);
        save(dir + "/gh.bin", ol::dev2cpu(gH, (size_t)4 * n * n).get(), (size_t)4 * n * n);
        ol::gFresnelResponseBandLimit(gH, 2 * n, 2 * n, p, p, lambda, 0.002);
        save(dir + "/ghbl.bin", ol::dev2cpu(gH, (size_t)4 * n * n).get(), (size_t)4 * n * n);
        ol::AsmTransferFCheck(n, n, p, p, lambda, 0.05);
        ol::FresnelResponseCheck(n, n, p, p, lambda, 0.05);
        ol::shiftedFresnelPropInfo(u, n, n, p, p, lambda, 0.05, 0.5, 0.0, 0.0);
        ol::synchronize();
    } catch (const std::exception& e) {
        std::cout << "error: " << e.what() << "\n";
        return 1;
    }
    // _CghFresnelDiff (include/olcgh.h:315-349): the cloud in pixel units
    try {
        ol::Object<double> pix(m);
        for (int j = 0; j < m; j++)
            pix[j] = ol::Vect3<double>(xyz[3 * j] / p, xyz[3 * j + 1] / p, xyz[3 * j + 2] / p);
        auto ph = std::make_unique<float[]>((size_t)n * n);
        ol::_CghFresnelDiff(pix, ph, n, n, p, lambda);
        save(dir + "/cgh_diff_phase.bin", ph.get(), (size_t)n * n);
    } catch (const std::exception& e) {
        std::cout << "error: " << e.what() << "\n";
        return 1;
    }
    // tests/Prop_test.cpp:46 unchanged, with its tests/para.h:8-18 parameters: the
    // reference's 3840 x 2160 demo field (padded 7680 x 4320, mixed radix)
    {
        std::ifstream probe(dir + "/field4k.bin");
        if (probe) {
            typedef float PREC_T;
            const int WIDTH = 1920 * 2;
            const int HEIGHT = 1080 * 2;
            PREC_T p = 3.74e-6;
            PREC_T lambda = 532e-9;
            PREC_T d = -0.1 + 0;
            const auto f4 = load<std::complex<float>>(dir + "/field4k.bin", (size_t)WIDTH * HEIGHT);
            std::unique_ptr<std::complex<PREC_T>[]> u;
            u = std::make_unique<std::complex<PREC_T>[]>(WIDTH * HEIGHT);
            for (size_t i = 0; i < (size_t)WIDTH * HEIGHT; i++) u[i] = f4[i];
            try {
                ol::FresnelProp(u,HEIGHT,WIDTH,p,p,lambda,-d);
            } catch (const std::exception& e) {
                std::cout << "error: " << e.what() << "\n";
                return 1;
            }
            save(dir + "/fresnel4k.bin", u.get(), (size_t)WIDTH * HEIGHT);
        }
    }
    // errors are exceptions, not exit(1): a side whose padded length 2 x 17 has a
    // prime factor above 13 is refused
    try {
        auto bad = std::make_unique<std::complex<float>[]>(17 * 17);
        ol::AsmProp(bad, 17, 17, p, p, lambda, 0.05);
        std::cout << "no exception for a 17x17 field\n";
        return 1;
    } catch (const std::runtime_error& e) {
        std::cout << "caught: " << e.what() << "\n";
    }
    std::cout << "done\n";
    return 0;
}
