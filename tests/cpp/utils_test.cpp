// Exercises the array helpers of include/olutils.h (host and device overloads) and
// the random lenslet forms of include/Lensarry.h the way the reference's callers use
// them, writing each result to <dir>/<name>.bin for the Python test to check.
//   utils_test <dir>   reads <dir>/a.bin, <dir>/b.bin (complex64 ny x nx) and
//                      <dir>/params.txt ("ny nx")
#include <cstdint>
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
static std::unique_ptr<T[]> load(const std::string& path, size_t count) {
    auto v = std::make_unique<T[]>(count);
    std::ifstream f(path, std::ios::binary);
    f.read(reinterpret_cast<char*>(v.get()), sizeof(T) * count);
    if (!f) throw std::runtime_error("short read: " + path);
    return v;
}
template <typename T>
static std::unique_ptr<T[]> copy(const std::unique_ptr<T[]>& src, size_t n) {
    auto v = std::make_unique<T[]>(n);
    for (size_t i = 0; i < n; i++) v[i] = src[i];
    return v;
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string dir = argv[1];
    int64_t ny = 0, nx = 0;
    {
        std::ifstream f(dir + "/params.txt");
        f >> ny >> nx;
    }
    const size_t n = (size_t)(ny * nx);
    typedef std::complex<float> C;
    try {
        auto a = load<C>(dir + "/a.bin", n);
        auto b = load<C>(dir + "/b.bin", n);
        // host FFT pair, out of place and in place
        std::unique_ptr<C[]> F;
        ol::fft(a, F, ny, nx);
        save(dir + "/fft.bin", F.get(), n);
        auto G = copy(a, n);
        ol::ifft(G, G, ny, nx);
        save(dir + "/ifft.bin", G.get(), n);
        // complex<double> data through the same transform
        auto ad = std::make_unique<std::complex<double>[]>(n);
        for (size_t i = 0; i < n; i++) ad[i] = std::complex<double>(a[i]);
        ol::fft(ad, ad, ny, nx);
        save(dir + "/fft_c128.bin", ad.get(), n);
        // elementwise helpers
        auto m = copy(a, n);
        ol::mul_complex(m, b, m, ny, nx);
        save(dir + "/mul.bin", m.get(), n);
        auto md = std::make_unique<std::complex<double>[]>(n);
        auto bd = std::make_unique<std::complex<double>[]>(n);
        for (size_t i = 0; i < n; i++) {
            md[i] = std::complex<double>(a[i]);
            bd[i] = std::complex<double>(b[i]);
        }
        ol::mul_complex(md, bd, md, ny, nx);
        save(dir + "/mul_c128.bin", md.get(), n);
        auto s = copy(a, n);
        ol::mul_scalar(s, 0.37, ny, nx);
        save(dir + "/scalar.bin", s.get(), n);
        auto st = copy(a, n);
        ol::standardization(st, ny, nx);
        save(dir + "/standard.bin", st.get(), n);
        auto sh = copy(a, n);
        ol::fftshift(sh, ny, nx);
        save(dir + "/shift.bin", sh.get(), n);
        // padding, cropping, cutting
        std::unique_ptr<C[]> p2;
        ol::zeropadding(a, p2, ny, nx);
        save(dir + "/pad2.bin", p2.get(), (size_t)(4 * n));
        auto pg = copy(a, n);
        ol::zeropadding(pg, pg, ny, nx, ny + 7, nx + 4);
        save(dir + "/padg.bin", pg.get(), (size_t)((ny + 7) * (nx + 4)));
        ol::del_zero(p2, p2, 2 * ny, 2 * nx);
        save(dir + "/delzero.bin", p2.get(), n);
        std::unique_ptr<C[]> c;
        ol::cut(a, c, ny, nx, 3, 5, ny / 2, nx / 3);
        save(dir + "/cut.bin", c.get(), (size_t)((ny / 2) * (nx / 3)));
        auto img = std::make_unique<uint8_t[]>(n);
        for (size_t i = 0; i < n; i++) img[i] = (uint8_t)(i * 7 + 3);
        std::unique_ptr<uint8_t[]> imgp;
        ol::zeropadding(img, imgp, ny, nx, ny + 2, nx + 6);
        save(dir + "/padu8.bin", imgp.get(), (size_t)((ny + 2) * (nx + 6)));

        // device overloads
        auto da = ol::cpu2dev(a, n);
        auto db = ol::cpu2dev(b, n);
        ol::device_ptr<C[]> dF;
        dF = ol::make_device<C[]>(n);
        ol::gFFT plan(ny, nx);
        plan.fft(da, dF);
        save(dir + "/gfft.bin", ol::dev2cpu(dF, n).get(), n);
        plan.ifft(dF, dF);
        save(dir + "/gfft_back.bin", ol::dev2cpu(dF, n).get(), n);
        auto dm = ol::make_device<C[]>(n);
        ol::mult(da, db, dm, (int)ny, (int)nx);
        save(dir + "/gmult.bin", ol::dev2cpu(dm, n).get(), n);
        ol::add(da, db, dm, (int)ny, (int)nx);
        save(dir + "/gadd.bin", ol::dev2cpu(dm, n).get(), n);
        ol::mul_scalar(dm, 2.5f, (int)ny, (int)nx);
        save(dir + "/gscalar.bin", ol::dev2cpu(dm, n).get(), n);
        auto dsh = ol::cpu2dev(a, n);
        ol::gfftshift(dsh, (int)ny, (int)nx);
        save(dir + "/gshift.bin", ol::dev2cpu(dsh, n).get(), n);
        ol::device_ptr<C[]> dp;
        ol::gzeropadding(da, dp, (int)ny, (int)nx);
        save(dir + "/gpad2.bin", ol::dev2cpu(dp, 4 * n).get(), 4 * n);
        ol::gdel_zero(dp, dp, 2 * (int)ny, 2 * (int)nx);
        save(dir + "/gdelzero.bin", ol::dev2cpu(dp, n).get(), n);
        ol::device_ptr<C[]> dc;
        ol::gcut(da, dc, (int)ny, (int)nx, 3, 5, (int)ny / 2, (int)nx / 3);
        save(dir + "/gcut.bin", ol::dev2cpu(dc, (size_t)((ny / 2) * (nx / 3))).get(),
             (size_t)((ny / 2) * (nx / 3)));

        // random lenslet forms (include/Lensarry.h:163-328)
        const float pitch = 8e-6f, bw = 16 * 8e-6f, f = 0.01f, lambda = 532e-9f, pl = 60e-6f;
        ol::Lensarry<float> L((int)ny, (int)nx, pitch, bw, f, lambda);
        L.gen_complex_random(pl, ol::PLAY, 2, -3);
        save(dir + "/lens_random.bin", L.u.get(), n);
        L.gen_complex_limit_random(pl, ol::RECORD, 0, 0);
        save(dir + "/lens_limit_random.bin", L.u.get(), n);
        ol::Lensarry<float> L2((int)ny, (int)nx, pitch, bw, f, lambda);
        auto ul = copy(a, n);
        L2.mult_complex_random(ul, pl, ol::PLAY, 1, 1);
        save(dir + "/lens_mult_random.bin", ul.get(), n);
        ol::Lensarry<float> L3((int)ny, (int)nx, pitch, bw, f, lambda);
        auto um = copy(a, n);
        L3.mult_complex(um, ol::PLAY, 0, 0);
        save(dir + "/lens_mult.bin", um.get(), n);
        auto blk = std::make_unique<C[]>((size_t)L.blockpx * L.blockpx);
        for (int i = 0; i < L.blockpx * L.blockpx; i++) blk[i] = C((float)i, 1.0f - i);
        auto dup = L.duplicate_multrandom(blk, pl, 0, 0);
        save(dir + "/lens_dup_multrandom.bin", dup.get(), n);
        ol::synchronize();
    } catch (const std::exception& e) {
        std::cout << "error: " << e.what() << "\n";
        return 1;
    }
    // a window outside the array is an error, not undefined behaviour
    try {
        auto a = std::make_unique<C[]>(16);
        std::unique_ptr<C[]> c;
        ol::cut(a, c, 4, 4, 2, 2, 3, 3);
        std::cout << "no exception for a cut outside the array\n";
        return 1;
    } catch (const std::runtime_error& e) {
        std::cout << "caught: " << e.what() << "\n";
    }
    std::cout << "done\n";
    return 0;
}
