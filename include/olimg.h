// olimg.h -- drop-in for the 8-bit encodings and BMP I/O of the reference's
// include/olimg.h, over libolx.so.
//
// The encodings (phase8bit, quant8bit, complex2img, complex2real) run on the device:
// host-array overloads stage the input to HBM and the result back; the device_ptr
// overloads stay in HBM.  Results equal the reference's x86-64 build byte for byte
// (FP64 arithmetic with one rounding per operation, C round(), and the compiled
// double -> uint8 conversion, which wraps phase8bit's negative values modulo 256).
// BMP files are host I/O.  The OpenCV paths of the reference (imread / imwrite,
// cvMat2sptr, non-.bmp Save) are not provided: Save throws for other extensions.
#ifndef OLIMG_H
#define OLIMG_H
#include <complex>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <type_traits>

#include "oldevice.h"
#include "olx.h"

namespace ol {

// include/olimg.h:268
enum MODE { PHASE, REAL, IMAG, AMP };

namespace detail {

template <typename T>
constexpr uint32_t real_flag() {
    static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value,
                  "encodings take float or double data");
    return std::is_same<T, double>::value ? OLX_IN_F64 : 0u;
}

inline int img_mode(MODE m) {
    switch (m) {
        case REAL: return OLX_IMG_REAL;
        case IMAG: return OLX_IMG_IMAG;
        case AMP: return OLX_IMG_AMP;
        default: return OLX_IMG_PHASE;
    }
}

// n host values of T into staging slot 0
template <typename T>
void* stage_in(const T* p, std::size_t n) {
    void* d = staging(0).get(sizeof(T) * (n ? n : 1));
    if (n) check(olx_copy_h2d(ctx(), d, p, sizeof(T) * n), "stage");
    return d;
}

inline uint8_t* img_buf(std::size_t n) { return static_cast<uint8_t*>(staging(1).get(n ? n : 1)); }

inline void img_back(uint8_t* img, const void* d, std::size_t n) {
    if (n) check(olx_copy_d2h(ctx(), img, d, n), "copy image");
}

inline bool is_bmp(const char* path) {
    const char* ext = std::strrchr(path, '.');
    return ext && std::strcmp(ext, ".bmp") == 0;
}

}  // namespace detail

// ------------------------------------------------------------------ BMP files
// include/olimg.h:56-99
inline void bmpread(const char* fname, uint8_t* img, int ny, int nx) {
    detail::check(olx_bmp_read(fname, img, ny, nx), "bmpread");
}
// include/olimg.h:130-133 (the reference decodes through OpenCV first, then bmpread)
inline void bmpread(const char* fname, std::unique_ptr<uint8_t[]>& img, int ny, int nx) {
    if (!img) img = std::make_unique<uint8_t[]>((std::size_t)ny * nx);
    bmpread(fname, img.get(), ny, nx);
}
// include/olimg.h:136-165
inline void bmpwrite(const char* fname, uint8_t* img, int ny, int nx) {
    detail::check(olx_bmp_write(fname, img, ny, nx), "bmpwrite");
}
// include/olimg.h:167-169
inline void bmpwrite(const char* fname, std::unique_ptr<uint8_t[]>& img, int ny, int nx) {
    bmpwrite(fname, img.get(), ny, nx);
}

// ------------------------------------------------------------------ real -> 8 bit
// include/olimg.h:178-198
template <typename PREC_T>
void quant8bit(PREC_T* real, uint8_t* img, int height, int width) {
    const std::size_t n = (std::size_t)height * width;
    void* d = detail::stage_in(real, n);
    uint8_t* di = detail::img_buf(n);
    detail::check(olx_quant8bit(detail::ctx(), d, di, (int64_t)n, detail::real_flag<PREC_T>()),
                  "quant8bit");
    detail::img_back(img, di, n);
}
// include/olimg.h:200-203 (PREC_T deduces to T[] for unique_ptr<T[]>, as there)
template <typename PREC_T>
void quant8bit(std::unique_ptr<PREC_T>& real, std::unique_ptr<uint8_t[]>& img, int height, int width) {
    quant8bit(real.get(), img.get(), height, width);
}

// include/olimg.h:205-213
template <typename REAL_T>
void phase8bit(REAL_T* real, uint8_t* img, int height, int width) {
    const std::size_t n = (std::size_t)height * width;
    void* d = detail::stage_in(real, n);
    uint8_t* di = detail::img_buf(n);
    detail::check(olx_phase8bit(detail::ctx(), d, di, (int64_t)n, detail::real_flag<REAL_T>()),
                  "phase8bit");
    detail::img_back(img, di, n);
}
// include/olimg.h:215-218
template <typename PREC_T>
void phase8bit(std::unique_ptr<PREC_T>& real, std::unique_ptr<uint8_t[]>& img, int height, int width) {
    phase8bit(real.get(), img.get(), height, width);
}

// ------------------------------------------------------------------ complex -> 8 bit / real
namespace detail {
template <typename PREC_T>
void complex2img_mode(const std::complex<PREC_T>* u, uint8_t* img, std::size_t n, int mode,
                      bool on_device) {
    const void* d = on_device ? static_cast<const void*>(u) : stage_in(u, n);
    uint8_t* di = img_buf(n);
    check(olx_complex2img(ctx(), d, di, (int64_t)n, mode | (int)real_flag<PREC_T>()), "complex2img");
    img_back(img, di, n);
}
}  // namespace detail

// include/olimg.h:314-346: realmode picks Re u (maximum seeded with 0), else |u|
template <typename PREC_T>
void complex2img(std::unique_ptr<std::complex<PREC_T>[]>& u, std::unique_ptr<uint8_t[]>& img,
                 int ny, int nx, bool realmode = false) {
    detail::complex2img_mode(u.get(), img.get(), (std::size_t)ny * nx,
                             realmode ? OLX_IMG_REAL_LEGACY : OLX_IMG_AMP, false);
}
// include/olimg.h:348-391
template <typename PREC_T>
void complex2img(std::unique_ptr<std::complex<PREC_T>[]>& u, std::unique_ptr<uint8_t[]>& img,
                 int ny, int nx, MODE mode) {
    detail::complex2img_mode(u.get(), img.get(), (std::size_t)ny * nx, detail::img_mode(mode), false);
}
// include/olimg.h:422-428
template <typename PREC_T>
std::unique_ptr<uint8_t[]> complex2img(std::unique_ptr<std::complex<PREC_T>[]>& u, int ny, int nx,
                                       MODE mode) {
    auto img = std::make_unique<uint8_t[]>((std::size_t)ny * nx);
    complex2img(u, img, ny, nx, mode);
    return img;
}
// device field (no reference counterpart: the reference encodes on the host)
template <typename PREC_T>
void complex2img(device_ptr<std::complex<PREC_T>[]>& u, std::unique_ptr<uint8_t[]>& img, int ny,
                 int nx, MODE mode) {
    if (!img) img = std::make_unique<uint8_t[]>((std::size_t)ny * nx);
    detail::complex2img_mode(u.get(), img.get(), (std::size_t)ny * nx, detail::img_mode(mode), true);
}

// include/olimg.h:393-413
template <typename PREC_T>
void complex2real(std::unique_ptr<std::complex<PREC_T>[]>& u, std::unique_ptr<PREC_T[]>& real,
                  int ny, int nx, MODE mode) {
    const std::size_t n = (std::size_t)ny * nx;
    const void* d = detail::stage_in(u.get(), n);
    void* dr = detail::staging(1).get(sizeof(PREC_T) * (n ? n : 1));
    detail::check(olx_complex2real(detail::ctx(), d, dr, (int64_t)n,
                                   detail::img_mode(mode) | (int)detail::real_flag<PREC_T>()),
                  "complex2real");
    if (n) detail::check(olx_copy_d2h(detail::ctx(), real.get(), dr, sizeof(PREC_T) * n), "complex2real");
}
// include/olimg.h:415-420
template <typename PREC_T>
std::unique_ptr<PREC_T[]> complex2real(std::unique_ptr<std::complex<PREC_T>[]>& u, int ny, int nx,
                                       MODE mode) {
    auto real = std::make_unique<PREC_T[]>((std::size_t)ny * nx);
    complex2real(u, real, ny, nx, mode);
    return real;
}

// ------------------------------------------------------------------ Save
// include/olimg.h:430-442 (.bmp only)
template <typename COMPLEX_T>
void Save(const char* path, std::unique_ptr<std::complex<COMPLEX_T>[]>& u, int height, int width,
          MODE savemode = MODE::PHASE) {
    if (!detail::is_bmp(path)) throw std::runtime_error("Save: only .bmp files are written");
    auto img = std::make_unique<uint8_t[]>((std::size_t)width * height);
    complex2img(u, img, height, width, savemode);
    bmpwrite(path, img, height, width);
}
// include/olimg.h:444-454
inline void Save(const char* path, std::unique_ptr<float[]>& real, int64_t ny, int64_t nx,
                 MODE savemode = MODE::PHASE) {
    (void)savemode;  // unused there too
    if (!detail::is_bmp(path)) throw std::runtime_error("Save: only .bmp files are written");
    auto img = std::make_unique<uint8_t[]>((std::size_t)(ny * nx));
    quant8bit(real.get(), img.get(), (int)ny, (int)nx);
    bmpwrite(path, img, (int)ny, (int)nx);
}
// include/olimg.h:456-463: a device field, encoded in HBM (only the bytes come back)
template <typename COMPLEX_T>
void Save(const char* path, device_ptr<std::complex<COMPLEX_T>[]>& u, int height, int width,
          MODE savemode = MODE::PHASE) {
    if (!detail::is_bmp(path)) throw std::runtime_error("Save: only .bmp files are written");
    std::unique_ptr<uint8_t[]> img;
    complex2img(u, img, height, width, savemode);
    bmpwrite(path, img, height, width);
}

}  // namespace ol
#endif
