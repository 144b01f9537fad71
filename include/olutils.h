// olutils.h -- drop-in for the reference's array helpers on MI355X: the host
// functions of include/olutils.h (std::unique_ptr arrays) and the device functions
// of include/golutils.h (ol::device_ptr arrays), same names, template parameters and
// argument order, all running as libolx.so kernels (include/olx.h).
//
// Host arrays are staged to HBM and the result is copied back.  Data movement
// (zeropadding, del_zero, cut, fftshift) and the elementwise complex128 products and
// scalings are exact for every element type; the 2-D FFT computes complex<float> data
// in complex64 and complex<double> data in complex128 (FP64 throughout, sides up to
// 8192).  Functions that return a new array (zeropadding,
// del_zero, cut and their device twins) move it into dst, which may be src, as the
// reference does.  Errors throw std::runtime_error.
#ifndef OLUTILS_H
#define OLUTILS_H
#include <complex>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "oldevice.h"
#include "olprop.h"

namespace ol {
namespace detail {

// a device copy of n host elements
template <typename T>
device_ptr<T[]> to_dev(const T* h, size_t n) {
    auto d = make_device<T[]>(n ? n : 1);
    if (n) check(olx_copy_h2d(ctx(), d.get(), h, sizeof(T) * n), "copy to device");
    return d;
}
template <typename T>
std::unique_ptr<T[]> to_host(const device_ptr<T[]>& d, size_t n) {
    auto h = std::make_unique<T[]>(n);
    if (n) check(olx_copy_d2h(ctx(), h.get(), d.get(), sizeof(T) * n), "copy to host");
    return h;
}

// device window copy: dst (dny x dnx) [h][w] = src[h - oy][w - ox], 0 outside
template <typename T>
void window(const void* src, int64_t sny, int64_t snx, void* dst, int64_t dny, int64_t dnx,
            int64_t oy, int64_t ox) {
    static_assert(sizeof(T) == 1 || sizeof(T) == 2 || sizeof(T) == 4 || sizeof(T) == 8 ||
                  sizeof(T) == 16, "element size must be 1, 2, 4, 8 or 16 bytes");
    check(olx_window_copy(ctx(), src, sny, snx, dst, dny, dnx, oy, ox, sizeof(T)), "window copy");
}

// host window copy through HBM into a new allocation moved into dst
template <typename T>
void host_window(std::unique_ptr<T[]>& src, std::unique_ptr<T[]>& dst, int64_t sny, int64_t snx,
                 int64_t dny, int64_t dnx, int64_t oy, int64_t ox) {
    auto ds = to_dev(src.get(), (size_t)(sny * snx));
    const size_t nd = (size_t)(dny * dnx);
    auto dd = make_device<T[]>(nd > 0 ? nd : 1);
    window<T>(ds.get(), sny, snx, dd.get(), dny, dnx, oy, ox);
    dst = to_host(dd, (size_t)(dny * dnx));
}

// device window copy into dst's buffer, or into a new one when dst is empty or src
template <typename T>
void dev_window(device_ptr<T[]>& src, device_ptr<T[]>& dst, int64_t sny, int64_t snx,
                int64_t dny, int64_t dnx, int64_t oy, int64_t ox) {
    if (src.get() == dst.get() || !dst) {
        const size_t nd = (size_t)(dny * dnx);
        auto tmp = make_device<T[]>(nd > 0 ? nd : 1);
        window<T>(src.get(), sny, snx, tmp.get(), dny, dnx, oy, ox);
        dst = std::move(tmp);
    } else {
        window<T>(src.get(), sny, snx, dst.get(), dny, dnx, oy, ox);
    }
}

template <typename T>
void fft_host(std::unique_ptr<std::complex<T>[]>& src, std::unique_ptr<std::complex<T>[]>& dst,
              int64_t ny, int64_t nx, int sign) {
    if (src.get() != dst.get()) {
        if (!dst) dst = std::make_unique<std::complex<T>[]>((size_t)(ny * nx));
        for (int64_t i = 0; i < ny * nx; i++) dst[i] = src[i];
    }
    if constexpr (std::is_same<T, double>::value)
        on_device_native(dst, ny * nx, [&](void* d) { check(olx_fft2d_c128(ctx(), d, ny, nx, sign), "fft"); });
    else
        on_device(dst, ny * nx, [&](void* d) { check(olx_fft2d_c64(ctx(), d, ny, nx, sign), "fft"); });
}

}  // namespace detail

// ------------------------------------------------------------------ host (olutils.h)
// include/olutils.h:19-41: unnormalised forward 2-D FFT, src -> dst (may alias)
template <typename PREC_T>
void fft(std::unique_ptr<std::complex<PREC_T>[]>& src, std::unique_ptr<std::complex<PREC_T>[]>& dst,
         int64_t ny, int64_t nx) {
    detail::fft_host(src, dst, ny, nx, -1);
}
// include/olutils.h:44-66: unnormalised backward 2-D FFT
template <typename PREC_T>
void ifft(std::unique_ptr<std::complex<PREC_T>[]>& src, std::unique_ptr<std::complex<PREC_T>[]>& dst,
          int64_t ny, int64_t nx) {
    detail::fft_host(src, dst, ny, nx, +1);
}

// include/olutils.h:81-89: u *= c (the scalar converted to the element type)
template <typename COMPLEX_T, typename PREC_T>
void mul_scalar(std::unique_ptr<std::complex<COMPLEX_T>[]>& u, PREC_T c, int64_t ny, int64_t nx) {
    static_assert(std::is_same<COMPLEX_T, float>::value || std::is_same<COMPLEX_T, double>::value,
                  "COMPLEX_T must be float or double");
    const size_t n = (size_t)(ny * nx);
    auto d = detail::to_dev(u.get(), n);
    if constexpr (std::is_same<COMPLEX_T, float>::value)
        detail::check(olx_scale_c64(detail::ctx(), d.get(), (int64_t)n, (float)c), "mul_scalar");
    else
        detail::check(olx_scale_c128(detail::ctx(), d.get(), (int64_t)n, (double)c), "mul_scalar");
    detail::check(olx_copy_d2h(detail::ctx(), u.get(), d.get(), sizeof(std::complex<COMPLEX_T>) * n),
                  "copy to host");
}

// include/olutils.h:70-79: u *= 1 / (ny nx), the factor rounded to float as there
template <typename PREC_T>
void standardization(std::unique_ptr<std::complex<PREC_T>[]>& u, int64_t ny, int64_t nx) {
    const float c = 1.0 / (ny * nx);
    mul_scalar(u, c, ny, nx);
}

// include/olutils.h:95-117: quadrant swap in place
template <typename PREC_T>
void fftshift(std::unique_ptr<std::complex<PREC_T>[]>& u, int64_t height, int64_t width) {
    const size_t n = (size_t)(height * width);
    auto d = detail::to_dev(u.get(), n);
    detail::check(olx_fftshift(detail::ctx(), d.get(), height, width, sizeof(std::complex<PREC_T>)),
                  "fftshift");
    detail::check(olx_copy_d2h(detail::ctx(), u.get(), d.get(), sizeof(std::complex<PREC_T>) * n),
                  "copy to host");
}

// include/olutils.h:121-130: out = in1 * in2 elementwise (out may be in1 or in2)
template <typename PREC_T>
void mul_complex(std::unique_ptr<std::complex<PREC_T>[]>& in1, std::unique_ptr<std::complex<PREC_T>[]>& in2,
                 std::unique_ptr<std::complex<PREC_T>[]>& out, size_t height, size_t width) {
    static_assert(std::is_same<PREC_T, float>::value || std::is_same<PREC_T, double>::value,
                  "PREC_T must be float or double");
    const size_t n = height * width;
    auto a = detail::to_dev(in1.get(), n);
    auto b = detail::to_dev(in2.get(), n);
    if constexpr (std::is_same<PREC_T, float>::value)
        detail::check(olx_mul_c64(detail::ctx(), a.get(), b.get(), a.get(), (int64_t)n), "mul_complex");
    else
        detail::check(olx_mul_c128(detail::ctx(), a.get(), b.get(), a.get(), (int64_t)n), "mul_complex");
    if (!out) out = std::make_unique<std::complex<PREC_T>[]>(n);
    detail::check(olx_copy_d2h(detail::ctx(), out.get(), a.get(), sizeof(std::complex<PREC_T>) * n),
                  "copy to host");
}

// include/olutils.h:133-153: centred 2x padding, out is 2 in_height x 2 in_width
template <typename PREC_T>
void zeropadding(std::unique_ptr<std::complex<PREC_T>[]>& in, std::unique_ptr<std::complex<PREC_T>[]>& out,
                 size_t in_height, size_t in_width) {
    const int64_t h = (int64_t)in_height, w = (int64_t)in_width;
    detail::host_window(in, out, h, w, 2 * h, 2 * w, h / 2, w / 2);
}
// include/olutils.h:155-178: centred padding to out_height x out_width (left pads
// (out - in) / 2), any element type of 1, 2, 4, 8 or 16 bytes
template <typename _Tp>
void zeropadding(std::unique_ptr<_Tp[]>& src, std::unique_ptr<_Tp[]>& dst, int64_t in_height,
                 int64_t in_width, int64_t out_height, int64_t out_width) {
    detail::host_window(src, dst, in_height, in_width, out_height, out_width,
                        (out_height - in_height) / 2, (out_width - in_width) / 2);
}
// include/olutils.h:181-198: keep the centre [h/4, 3h/4) x [w/4, 3w/4) of an
// in_height x in_width array (out is in_height/2 x in_width/2)
template <typename PREC_T>
void del_zero(std::unique_ptr<std::complex<PREC_T>[]>& in, std::unique_ptr<std::complex<PREC_T>[]>& out,
              int64_t in_height, int64_t in_width) {
    detail::host_window(in, out, in_height, in_width, in_height / 2, in_width / 2, -(in_height / 4),
                        -(in_width / 4));
}
// include/olutils.h:200-211: the ysize x xsize window at (y, x) of a height x width
// array (the window must lie inside it)
template <typename _Tp>
void cut(std::unique_ptr<_Tp[]>& src, std::unique_ptr<_Tp[]>& dst, int64_t height, int64_t width,
         int64_t y, int64_t x, int64_t ysize, int64_t xsize) {
    if (y < 0 || x < 0 || ysize < 0 || xsize < 0 || y + ysize > height || x + xsize > width)
        throw std::runtime_error("cut: window outside the array");
    detail::host_window(src, dst, height, width, ysize, xsize, -y, -x);
}

// ------------------------------------------------------------------ device (golutils.h)
// include/golutils.h:14-53: a 2-D transform of fixed size, reused across calls
class gFFT {
  public:
    int64_t ny = 0, nx = 0;
    bool flag = false;
    gFFT() {}
    gFFT(int64_t ny, int64_t nx) { set(ny, nx); }
    void set(int64_t ny_, int64_t nx_) {
        ny = ny_;
        nx = nx_;
        flag = true;
    }
    void fft(device_ptr<std::complex<float>[]>& src, device_ptr<std::complex<float>[]>& dst) { run(src, dst, -1); }
    void ifft(device_ptr<std::complex<float>[]>& src, device_ptr<std::complex<float>[]>& dst) { run(src, dst, 1); }

  private:
    void run(device_ptr<std::complex<float>[]>& src, device_ptr<std::complex<float>[]>& dst, int sign) {
        if (!flag) throw std::runtime_error("gFFT: plan not set");
        if (src.get() != dst.get())
            detail::window<std::complex<float>>(src.get(), ny, nx, dst.get(), ny, nx, 0, 0);
        detail::check(olx_fft2d_c64(detail::ctx(), dst.get(), ny, nx, sign), "gFFT");
    }
};

// include/golutils.h:55-76: unnormalised 2-D FFT of device arrays, src -> dst (may alias)
template <typename PREC_T>
void fft(device_ptr<std::complex<PREC_T>[]>& src, device_ptr<std::complex<PREC_T>[]>& dst, int64_t ny, int64_t nx) {
    static_assert(std::is_same<PREC_T, float>::value, "device fields are complex<float>");
    gFFT(ny, nx).fft(src, dst);
}
template <typename PREC_T>
void ifft(device_ptr<std::complex<PREC_T>[]>& src, device_ptr<std::complex<PREC_T>[]>& dst, int64_t ny, int64_t nx) {
    static_assert(std::is_same<PREC_T, float>::value, "device fields are complex<float>");
    gFFT(ny, nx).ifft(src, dst);
}

// include/golutils.h:80-110
template <typename PREC_T>
void gfftshift(device_ptr<std::complex<PREC_T>[]>& u, int ny, int nx) {
    detail::check(olx_fftshift(detail::ctx(), u.get(), ny, nx, sizeof(std::complex<PREC_T>)), "gfftshift");
}

// include/golutils.h:123-143: c = a * b
template <typename PREC_T>
void mult(device_ptr<std::complex<PREC_T>[]>& a, device_ptr<std::complex<PREC_T>[]>& b,
          device_ptr<std::complex<PREC_T>[]>& c, int height, int width) {
    static_assert(std::is_same<PREC_T, float>::value, "device fields are complex<float>");
    detail::check(olx_mul_c64(detail::ctx(), a.get(), b.get(), c.get(), (int64_t)height * width), "mult");
}
// include/golutils.h:145-164: udst = u1 + u2
template <typename PREC_T = float>
void add(device_ptr<std::complex<PREC_T>[]>& u1, device_ptr<std::complex<PREC_T>[]>& u2,
         device_ptr<std::complex<PREC_T>[]>& udst, int height, int width) {
    static_assert(std::is_same<PREC_T, float>::value, "device fields are complex<float>");
    detail::check(olx_add_c64(detail::ctx(), u1.get(), u2.get(), udst.get(), (int64_t)height * width), "add");
}
// include/golutils.h:166-187: u *= scalar
template <typename COMPLEX_T = float, typename PREC_T = float>
void mul_scalar(device_ptr<std::complex<COMPLEX_T>[]>& u, PREC_T scalar, int height, int width) {
    static_assert(std::is_same<COMPLEX_T, float>::value, "device fields are complex<float>");
    detail::check(olx_scale_c64(detail::ctx(), u.get(), (int64_t)height * width, (float)scalar), "mul_scalar");
}

// include/golutils.h:189-205: centred 2x padding (dst 2 in_height x 2 in_width)
template <typename PREC_T>
void gzeropadding(device_ptr<std::complex<PREC_T>[]>& src, device_ptr<std::complex<PREC_T>[]>& dst,
                  int in_height, int in_width) {
    detail::dev_window(src, dst, in_height, in_width, 2 * (int64_t)in_height, 2 * (int64_t)in_width,
                       in_height / 2, in_width / 2);
}
// include/golutils.h:207-244: centred padding to out_height x out_width
template <typename _Tp>
void gzeropadding(device_ptr<_Tp[]>& src, device_ptr<_Tp[]>& dst, int in_height, int in_width,
                  int out_height, int out_width) {
    detail::dev_window(src, dst, in_height, in_width, out_height, out_width,
                       (out_height - in_height) / 2, (out_width - in_width) / 2);
}
// include/golutils.h:247-278: keep the centre half of an ny x nx array
template <typename _Tp>
void gdel_zero(device_ptr<_Tp[]>& src, device_ptr<_Tp[]>& dst, int ny, int nx) {
    detail::dev_window(src, dst, ny, nx, ny / 2, nx / 2, -(ny / 4), -(nx / 4));
}
// include/golutils.h:281-308: the ysize x xsize window at (y, x)
template <typename _Tp>
void gcut(device_ptr<_Tp[]>& src, device_ptr<_Tp[]>& dst, int ny, int nx, int y, int x, int ysize,
          int xsize) {
    if (y < 0 || x < 0 || ysize < 0 || xsize < 0 || y + ysize > ny || x + xsize > nx)
        throw std::runtime_error("gcut: window outside the array");
    detail::dev_window(src, dst, ny, nx, ysize, xsize, -y, -x);
}

}  // namespace ol
#endif
