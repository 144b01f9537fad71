// olprop.h -- drop-in for the reference's propagation API on MI355X.
//
// Host overloads keep the signatures of include/olprop.h (namespace ol, same
// template parameters and argument order): the field stays in the caller's
// std::unique_ptr<std::complex<COMPLEX_T>[]> allocation, is staged to HBM, runs
// through the fused HIP kernels of libolx.so, and is copied back in place.
// COMPLEX_T = float fields are computed in complex64 with FP64 phase arithmetic
// (PREC_T=double semantics, unless OL_PREC_T_FLOAT_COMPAT asks for the reference's
// float arithmetic where PREC_T = float).  COMPLEX_T = double fields of AsmProp,
// mAsmProp, FresnelProp, FresnelPropBandLimit and Prop stay complex128 and run the
// FP64 passes (sides up to 4096: larger ones throw); the other complex<double> calls
// are rounded to complex64.
// Device overloads mirror include/golprop.h on ol::device_ptr<std::complex<float>[]>.
// Errors throw std::runtime_error; nothing is printed and nothing exits.
#ifndef OLPROP_H
#define OLPROP_H
#include <climits>
#include <cmath>
#include <complex>
#include <cstdint>
#include <iomanip>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "oldefine.h"
#include "oldevice.h"

namespace ol {
namespace detail {

// stage a host field to HBM, run fn(device_ptr), copy back into the same allocation
template <typename C, typename F>
void on_device(std::unique_ptr<std::complex<C>[]>& u, int64_t count, F&& fn) {
    static_assert(std::is_same<C, float>::value || std::is_same<C, double>::value,
                  "COMPLEX_T must be float or double");
    const size_t bytes = sizeof(std::complex<float>) * (size_t)count;
    void* d = staging(0).get(bytes);
    if constexpr (std::is_same<C, float>::value) {
        check(olx_copy_h2d(ctx(), d, u.get(), bytes), "copy to device");
        fn(d);
        check(olx_copy_d2h(ctx(), u.get(), d, bytes), "copy to host");
    } else {
        std::vector<std::complex<float>> t((size_t)count);
        for (int64_t i = 0; i < count; i++) t[i] = std::complex<float>(u[i]);
        check(olx_copy_h2d(ctx(), d, t.data(), bytes), "copy to device");
        fn(d);
        check(olx_copy_d2h(ctx(), t.data(), d, bytes), "copy to host");
        for (int64_t i = 0; i < count; i++) u[i] = std::complex<C>(t[i]);
    }
}

// stage a host field to HBM in its own precision, run fn(device pointer), copy back
template <typename C, typename F>
void on_device_native(std::unique_ptr<std::complex<C>[]>& u, int64_t count, F&& fn) {
    const size_t bytes = sizeof(std::complex<C>) * (size_t)count;
    void* d = staging(0).get(bytes);
    check(olx_copy_h2d(ctx(), d, u.get(), bytes), "copy to device");
    fn(d);
    check(olx_copy_d2h(ctx(), u.get(), d, bytes), "copy to host");
}

inline int64_t win(double v) {
    const double f = std::floor(v);
    if (!(f == f) || f >= (double)INT_MAX) return INT_MAX;
    if (f <= (double)INT_MIN) return INT_MIN;
    return (int64_t)f;
}

}  // namespace detail

// ------------------------------------------------------------------ checks
// include/olprop.h:209-225
template <typename PREC_T = float>
bool AsmPropCheck(int64_t ny, int64_t nx, PREC_T dy, PREC_T dx, PREC_T lambda, PREC_T z) {
    return olx_asm_prop_check(ny, nx, dy, dx, lambda, z) != 0;
}
// include/olprop.h:109-123
template <typename PREC_T = double>
bool FresnelPropCheck(size_t ny, size_t nx, PREC_T dy, PREC_T dx, PREC_T lambda, PREC_T z) {
    return olx_fresnel_prop_check((int64_t)ny, (int64_t)nx, dy, dx, lambda, z) != 0;
}

// ------------------------------------------------------------------ host API
// include/olprop.h:227-248 (net scale 1).  Defining OL_PREC_T_FLOAT_COMPAT before
// including the headers makes PREC_T = float calls use the reference's float
// arithmetic for the transfer function (OLX_PREC_F32); otherwise every PREC_T gets
// the FP64 phase.
namespace detail {
template <typename PREC_T>
constexpr uint32_t prec_flags() {
#ifdef OL_PREC_T_FLOAT_COMPAT
    return std::is_same<PREC_T, float>::value ? OLX_PREC_F32 : 0u;
#else
    return 0u;
#endif
}
}  // namespace detail
template <typename PREC_T = double, typename COMPLEX_T>
void AsmProp(std::unique_ptr<std::complex<COMPLEX_T>[]>& u, int64_t height, int64_t width,
             PREC_T dy, PREC_T dx, PREC_T lambda, PREC_T z) {
    if constexpr (std::is_same<COMPLEX_T, double>::value) {
        if (detail::prec_flags<PREC_T>() == 0u) {  // the fftw3 path: complex128 in FP64
            detail::on_device_native(u, height * width, [&](void* d) {
                detail::check(olx_asm_c128(detail::ctx(), d, height, width, dy, dx, lambda, z, 0),
                              "AsmProp");
            });
            return;
        }
    }
    detail::on_device(u, height * width, [&](void* d) {
        detail::check(olx_asm_c64(detail::ctx(), d, height, width, dy, dx, lambda, z,
                                  detail::prec_flags<PREC_T>()), "AsmProp");
    });
}

// include/molprop.h:80-101: the angular spectrum propagation the reference stages
// through host memory in blocks (fft2d_hyblid); net scale 1/M^2 as there
template <typename PREC_T = double, typename COMPLEX_T>
void mAsmProp(std::unique_ptr<std::complex<COMPLEX_T>[]>& u, int64_t height, int64_t width,
              PREC_T dy, PREC_T dx, PREC_T lambda, PREC_T z) {
    if constexpr (std::is_same<COMPLEX_T, double>::value) {
        detail::on_device_native(u, height * width, [&](void* d) {
            detail::check(olx_asm_c128(detail::ctx(), d, height, width, dy, dx, lambda, z,
                                       OLX_EXTRA_INV_M2), "mAsmProp");
        });
        return;
    }
    detail::on_device(u, height * width, [&](void* d) {
        detail::check(olx_asm_c64(detail::ctx(), d, height, width, dy, dx, lambda, z,
                                  OLX_EXTRA_INV_M2), "mAsmProp");
    });
}

// include/molprop.h:16-78: unnormalised 2-D FFT of a host array (forward / inverse),
// src -> dst (dst allocated when empty, may be src).  The reference moves blocks of
// blocknum_y rows / blocknum_x columns through device memory; the whole array fits
// in HBM here (sides: powers of two in [8, 16384]), so the block counts only keep
// the signature.
namespace detail {
template <typename PREC_T>
void hybrid_fft(std::unique_ptr<std::complex<PREC_T>[]>& src,
                std::unique_ptr<std::complex<PREC_T>[]>& dst, int64_t ny, int64_t nx, int sign) {
    static_assert(std::is_same<PREC_T, float>::value, "fft2d_hyblid: complex<float> data");
    if (&src != &dst) {
        if (!dst) dst = std::make_unique<std::complex<PREC_T>[]>((size_t)(ny * nx));
        for (int64_t i = 0; i < ny * nx; i++) dst[i] = src[i];
    }
    on_device(dst, ny * nx, [&](void* d) { check(olx_fft2d_c64(ctx(), d, ny, nx, sign), "fft2d_hyblid"); });
}
}  // namespace detail
template <typename PREC_T>
void fft2d_hyblid(std::unique_ptr<std::complex<PREC_T>[]>& src,
                  std::unique_ptr<std::complex<PREC_T>[]>& dst, int64_t ny, int64_t nx,
                  int blocknum_y = 2, int blocknum_x = 2) {
    (void)blocknum_y;
    (void)blocknum_x;
    detail::hybrid_fft(src, dst, ny, nx, -1);
}
template <typename PREC_T>
void ifft2d_hyblid(std::unique_ptr<std::complex<PREC_T>[]>& src,
                   std::unique_ptr<std::complex<PREC_T>[]>& dst, int64_t ny, int64_t nx,
                   int blocknum_y = 2, int blocknum_x = 2) {
    (void)blocknum_y;
    (void)blocknum_x;
    detail::hybrid_fft(src, dst, ny, nx, +1);
}

// include/olprop.h:125-148 (circconv / M^2)
template <typename PREC_T = double, typename COMPLEX_T>
void FresnelProp(std::unique_ptr<std::complex<COMPLEX_T>[]>& u, size_t height, size_t width,
                 PREC_T dy, PREC_T dx, PREC_T lambda, PREC_T z) {
    if constexpr (std::is_same<COMPLEX_T, double>::value) {
        detail::on_device_native(u, (int64_t)(height * width), [&](void* d) {
            detail::check(olx_fresnel_c128(detail::ctx(), d, (int64_t)height, (int64_t)width, dy,
                                           dx, lambda, z, 0), "FresnelProp");
        });
        return;
    }
    detail::on_device(u, (int64_t)(height * width), [&](void* d) {
        detail::check(olx_fresnel_c64(detail::ctx(), d, (int64_t)height, (int64_t)width, dy, dx,
                                      lambda, z, 0), "FresnelProp");
    });
}

// include/olprop.h:150-170 (per-axis window)
template <typename PREC_T = double, typename COMPLEX_T>
void FresnelPropBandLimit(std::unique_ptr<std::complex<COMPLEX_T>[]>& u, size_t height,
                          size_t width, PREC_T dy, PREC_T dx, PREC_T lambda, PREC_T z) {
    if constexpr (std::is_same<COMPLEX_T, double>::value) {
        detail::on_device_native(u, (int64_t)(height * width), [&](void* d) {
            detail::check(olx_fresnel_c128(detail::ctx(), d, (int64_t)height, (int64_t)width, dy,
                                           dx, lambda, z, OLX_FRESNEL_BL_CPU),
                          "FresnelPropBandLimit");
        });
        return;
    }
    detail::on_device(u, (int64_t)(height * width), [&](void* d) {
        detail::check(olx_fresnel_c64(detail::ctx(), d, (int64_t)height, (int64_t)width, dy, dx,
                                      lambda, z, OLX_FRESNEL_BL_CPU), "FresnelPropBandLimit");
    });
}

// include/olprop.h:338-408: CPU ARSS analogue (no windows, circconv / M^4), in place
template <typename PREC_T = float, typename COMPLEX_T>
void shiftedFresnelProp(std::unique_ptr<std::complex<COMPLEX_T>[]>& u, int64_t height,
                        int64_t width, PREC_T dy, PREC_T dx, PREC_T lambda, PREC_T z, PREC_T s,
                        PREC_T oy, PREC_T ox) {
    detail::on_device(u, height * width, [&](void* d) {
        detail::check(olx_arss_c64(detail::ctx(), d, d, height, width, dy, dx, lambda, z, s, oy,
                                   ox, OLX_ARSS_CPU), "shiftedFresnelProp");
    });
}

// include/olprop.h:307-330: shifted (off-axis) ASM, usrc -> udst (allocated when
// empty, may be usrc), with the shifted band-limit window of :250-305 and the
// reference's net scale 1/M^2
template <typename PREC_T = double, typename COMPLEX_T>
void shiftedAsmProp(std::unique_ptr<std::complex<COMPLEX_T>[]>& usrc,
                    std::unique_ptr<std::complex<COMPLEX_T>[]>& udst, int64_t height,
                    int64_t width, PREC_T dy, PREC_T dx, PREC_T lambda, PREC_T z, PREC_T oy,
                    PREC_T ox) {
    if (&usrc != &udst) {
        if (!udst) udst = std::make_unique<std::complex<COMPLEX_T>[]>((size_t)(height * width));
        for (int64_t i = 0; i < height * width; i++) udst[i] = usrc[i];
    }
    detail::on_device(udst, height * width, [&](void* d) {
        detail::check(olx_shifted_asm_c64(detail::ctx(), d, d, height, width, dy, dx, lambda, z,
                                          oy, ox, OLX_SHIFTED_CPU), "shiftedAsmProp");
    });
}

// include/olprop.h:614-651: blocknum_y x blocknum_x blocks, every block propagated to
// every block with shiftedAsmProp and summed, in place
template <typename PREC_T = float, typename COMPLEX_T>
void splitAsmProp(std::unique_ptr<std::complex<COMPLEX_T>[]>& u, int64_t height, int64_t width,
                  PREC_T dy, PREC_T dx, PREC_T lambda, PREC_T z, int blocknum_y = 2,
                  int blocknum_x = 2) {
    detail::on_device(u, height * width, [&](void* d) {
        detail::check(olx_split_prop_c64(detail::ctx(), d, height, width, dy, dx, lambda, z,
                                         OLX_MODE_ASM, blocknum_y, blocknum_x, OLX_SHIFTED_CPU),
                      "splitAsmProp");
    });
}

// include/golprop.h:773-813: the GPU split propagation of a host field (blocks through
// gshiftedAsmProp, or ARSS with s = 1 for FRESNEL), in place
template <typename PREC_T = float, typename COMPLEX_T>
void gsplitProp(std::unique_ptr<std::complex<COMPLEX_T>[]>& u, int64_t height, int64_t width,
                PREC_T dy, PREC_T dx, PREC_T lambda, PREC_T z, PROPMODE propmode,
                int blocknum_y = 2, int blocknum_x = 2) {
    if (propmode != ASM && propmode != FRESNEL)
        throw std::runtime_error("gsplitProp: propmode must be ASM or FRESNEL");
    detail::on_device(u, height * width, [&](void* d) {
        detail::check(olx_split_prop_c64(detail::ctx(), d, height, width, dy, dx, lambda, z,
                                         propmode == ASM ? OLX_MODE_ASM : OLX_MODE_FRESNEL,
                                         blocknum_y, blocknum_x, 0), "gsplitProp");
    });
}

// include/olprop.h:410-412
enum INTERPOLATION { BILINEAR, BICUBIC };

// include/olprop.h:414-526: from a source plane tilted by angle_src2dst to a parallel
// plane at distance z; in place, net scale 1/M^2 as the reference's
template <typename PREC_T = float, typename COMPLEX_T>
void TiltedSourceProp(std::unique_ptr<std::complex<COMPLEX_T>[]>& u, int ny, int nx, PREC_T dy,
                      PREC_T dx, PREC_T lambda, PREC_T z, PREC_T angle_src2dst,
                      INTERPOLATION inp = BILINEAR, PROPMODE propmode = ASM) {
    if (propmode != ASM && propmode != FRESNEL)
        throw std::runtime_error("TiltedSourceProp: propmode must be ASM or FRESNEL");
    detail::on_device(u, (int64_t)ny * nx, [&](void* d) {
        detail::check(olx_tilted_c64(detail::ctx(), d, ny, nx, dy, dx, lambda, z, angle_src2dst,
                                     inp == BICUBIC ? OLX_INTERP_BICUBIC : OLX_INTERP_BILINEAR,
                                     propmode == ASM ? OLX_MODE_ASM : OLX_MODE_FRESNEL,
                                     OLX_TILT_SOURCE), "TiltedSourceProp");
    });
}

// include/olprop.h:528-594: propagation by z, then the reference plane rotated by rotate
template <typename PREC_T = float, typename COMPLEX_T>
void TiltedReferenceProp(std::unique_ptr<std::complex<COMPLEX_T>[]>& u, int64_t ny, int64_t nx,
                         PREC_T dy, PREC_T dx, PREC_T lambda, PREC_T z, PREC_T rotate,
                         INTERPOLATION inp = BILINEAR, PROPMODE propmode = ASM) {
    if (propmode != ASM && propmode != FRESNEL)
        throw std::runtime_error("TiltedReferenceProp: propmode must be ASM or FRESNEL");
    detail::on_device(u, ny * nx, [&](void* d) {
        detail::check(olx_tilted_c64(detail::ctx(), d, ny, nx, dy, dx, lambda, z, rotate,
                                     inp == BICUBIC ? OLX_INTERP_BICUBIC : OLX_INTERP_BILINEAR,
                                     propmode == ASM ? OLX_MODE_ASM : OLX_MODE_FRESNEL,
                                     OLX_TILT_REFERENCE), "TiltedReferenceProp");
    });
}

// include/olprop.h:656-682 (silent; the reference prints the chosen mode)
template <typename COMPLEX_T, typename PREC_T>
void Prop(std::unique_ptr<std::complex<COMPLEX_T>[]>& u, int ny, int nx, PREC_T dy, PREC_T dx,
          PREC_T lambda, PREC_T d, PROPMODE propmode = PROPMODE::AUTO) {
    const int mode = propmode == ASM ? OLX_MODE_ASM : propmode == FRESNEL ? OLX_MODE_FRESNEL
                                                                          : OLX_MODE_AUTO;
    if constexpr (std::is_same<COMPLEX_T, double>::value) {
        detail::on_device_native(u, (int64_t)ny * nx, [&](void* dd) {
            detail::check(olx_prop_c128(detail::ctx(), dd, ny, nx, dy, dx, lambda, d, mode, nullptr),
                          "Prop");
        });
        return;
    }
    detail::on_device(u, (int64_t)ny * nx, [&](void* dd) {
        detail::check(olx_prop_c64(detail::ctx(), dd, ny, nx, dy, dx, lambda, d, mode, nullptr),
                      "Prop");
    });
}

// include/olprop.h:172-189 (+ band-limit window of :250-305 when bandlimit is set):
// the centred transfer function; evanescent entries are 0 (the reference: NaN)
template <typename PREC_T = float, typename COMPLEX_T>
void AsmTransferF(std::unique_ptr<std::complex<COMPLEX_T>[]>& H, int64_t ny, int64_t nx,
                  PREC_T dv, PREC_T du, PREC_T lambda, PREC_T z, bool bandlimit = false) {
    detail::on_device(H, ny * nx, [&](void* d) {
        detail::check(olx_asm_transfer_c64(detail::ctx(), d, ny, nx, dv, du, lambda, z,
                                           bandlimit ? OLX_BANDLIMIT : 0u), "AsmTransferF");
    });
}

// include/olprop.h:8-24
template <typename PREC_T = double, typename COMPLEX_T>
void FresnelResponse(std::unique_ptr<std::complex<COMPLEX_T>[]>& h, size_t height, size_t width,
                     PREC_T dy, PREC_T dx, PREC_T lambda, PREC_T z) {
    detail::on_device(h, (int64_t)(height * width), [&](void* d) {
        detail::check(olx_fresnel_response_c64(detail::ctx(), d, (int64_t)height, (int64_t)width,
                                               dy, dx, lambda, z, INT_MAX, INT_MAX),
                      "FresnelResponse");
    });
}

// include/olprop.h:26-54 (per-axis limits) and :56-84 (limits from p_limit)
template <typename PREC_T = double, typename COMPLEX_T>
void FresnelResponseBandLimit(std::unique_ptr<std::complex<COMPLEX_T>[]>& h, size_t height,
                              size_t width, PREC_T dy, PREC_T dx, PREC_T lambda, PREC_T z) {
    const int64_t wlim = detail::win(std::fabs(lambda * z * 0.5 / dx) / dx);
    const int64_t hlim = detail::win(std::fabs(lambda * z * 0.5 / dy) / dy);
    detail::on_device(h, (int64_t)(height * width), [&](void* d) {
        detail::check(olx_fresnel_response_c64(detail::ctx(), d, (int64_t)height, (int64_t)width,
                                               dy, dx, lambda, z, wlim, hlim),
                      "FresnelResponseBandLimit");
    });
}
template <typename PREC_T = double, typename COMPLEX_T>
void FresnelResponseBandLimit(std::unique_ptr<std::complex<COMPLEX_T>[]>& h, size_t height,
                              size_t width, PREC_T dy, PREC_T dx, PREC_T lambda, PREC_T z,
                              PREC_T p_limit) {
    const int64_t wlim = detail::win(std::fabs(lambda * z * 0.5 / p_limit) / dx);
    const int64_t hlim = detail::win(std::fabs(lambda * z * 0.5 / p_limit) / dy);
    detail::on_device(h, (int64_t)(height * width), [&](void* d) {
        detail::check(olx_fresnel_response_c64(detail::ctx(), d, (int64_t)height, (int64_t)width,
                                               dy, dx, lambda, z, wlim, hlim),
                      "FresnelResponseBandLimit");
    });
}

// include/olprop.h:250-305: the shifted transfer function exp(2 pi i (w z + ox u + oy v)),
// centred, with the off-axis window (computed in FP64; evanescent entries are 0)
template <typename PREC_T = float, typename COMPLEX_T>
void shiftedAsmTransferF(std::unique_ptr<std::complex<COMPLEX_T>[]>& H, int64_t ny, int64_t nx,
                         PREC_T dv, PREC_T du, PREC_T lambda, PREC_T z, PREC_T oy, PREC_T ox) {
    detail::on_device(H, ny * nx, [&](void* d) {
        detail::check(olx_shifted_asm_transfer_c64(detail::ctx(), d, ny, nx, dv, du, lambda, z, oy,
                                                   ox, OLX_BANDLIMIT),
                      "shiftedAsmTransferF");
    });
}

// include/olprop.h:87-105: prints the Fresnel regime's limits for the 2x padded field
// (shortest distance from each axis, the band-limit half widths and their pixel counts)
template <typename PREC_T = double>
void FresnelResponseCheck(size_t height, size_t width, PREC_T dy, PREC_T dx, PREC_T lambda,
                          PREC_T z) {
    const PREC_T half_w = dx * (PREC_T)(2 * width) / 2, half_h = dy * (PREC_T)(2 * height) / 2;
    const PREC_T xl = lambda * z * 0.5 / dx, yl = lambda * z * 0.5 / dy;
    std::cout << "zmin (from x) = " << std::setprecision(4) << 2 * half_w * dx / lambda << std::endl;
    std::cout << "zmin (from y) = " << std::setprecision(4) << 2 * half_h * dy / lambda << std::endl;
    std::cout << "xlim = " << std::setprecision(4) << xl << std::endl;
    std::cout << "ylim = " << std::setprecision(4) << yl << std::endl;
    std::cout << "widthlim[px] = " << (int)std::floor(xl / dx) * 2 << std::endl;
    std::cout << "heightlim[px] = " << (int)std::floor(yl / dy) * 2 << std::endl;
}

// include/olprop.h:191-208: prints the largest distance the ASM transfer function
// samples without aliasing, from each axis of the 2x padded field
template <typename PREC_T = double>
void AsmTransferFCheck(int64_t height, int64_t width, PREC_T dy, PREC_T dx, PREC_T lambda,
                       PREC_T z) {
    (void)z;
    const PREC_T du = 1 / (dx * (PREC_T)(2 * width)), dv = 1 / (dy * (PREC_T)(2 * height));
    const PREC_T umax = du * (PREC_T)width, vmax = dv * (PREC_T)height;
    const PREC_T root = std::sqrt(1 / (lambda * lambda) - umax * umax - vmax * vmax);
    std::cout << "zmax (from x) = " << std::setprecision(4) << root / (2 * umax * du) << std::endl;
    std::cout << "zmax (from y) = " << std::setprecision(4) << root / (2 * vmax * dv) << std::endl;
}

// include/olprop.h:332-336: an empty function in the reference
template <typename PREC_T = float, typename COMPLEX_T>
void shiftedFresnelPropInfo(std::unique_ptr<std::complex<COMPLEX_T>[]>&, int, int, PREC_T, PREC_T,
                            PREC_T, PREC_T, PREC_T, PREC_T, PREC_T) {}

// ------------------------------------------------------------------ device API
// (include/golprop.h; fields are ol::device_ptr<std::complex<float>[]>, updated in place)
using dfield = device_ptr<std::complex<float>[]>;

// include/golprop.h:325-332: the centred transfer function table in HBM
template <typename PREC_T>
void gAsmTransferF(dfield& H, int ny, int nx, PREC_T dv, PREC_T du, PREC_T lambda, PREC_T z) {
    detail::check(olx_asm_transfer_c64(detail::ctx(), H.get(), ny, nx, dv, du, lambda, z, 0),
                  "gAsmTransferF");
}
// include/golprop.h:361-419 (the reference's window is commented out: none here either)
template <typename PREC_T>
void gshiftedAsmTransferF(dfield& H, int ny, int nx, PREC_T dv, PREC_T du, PREC_T lambda,
                          PREC_T z, PREC_T oy, PREC_T ox) {
    detail::check(olx_shifted_asm_transfer_c64(detail::ctx(), H.get(), ny, nx, dv, du, lambda, z,
                                               oy, ox, 0),
                  "gshiftedAsmTransferF");
}
// include/golprop.h:58-66: the centred Fresnel impulse response in HBM
template <typename PREC_T>
void gFresnelResponse(dfield& H, int ny, int nx, PREC_T dy, PREC_T dx, PREC_T lambda, PREC_T z) {
    detail::check(olx_fresnel_response_c64(detail::ctx(), H.get(), ny, nx, dy, dx, lambda, z,
                                           INT_MAX, INT_MAX),
                  "gFresnelResponse");
}
// include/golprop.h:223-256: windowed by the x-limit on both axes
template <typename PREC_T>
void gFresnelResponseBandLimit(dfield& h, size_t ny, size_t nx, PREC_T dy, PREC_T dx,
                               PREC_T lambda, PREC_T z) {
    const int64_t lim = detail::win(std::fabs(lambda * z * 0.5 / dx) / dx);
    detail::check(olx_fresnel_response_c64(detail::ctx(), h.get(), (int64_t)ny, (int64_t)nx, dy,
                                           dx, lambda, z, lim, lim),
                  "gFresnelResponseBandLimit");
}
template <typename PREC_T>
void gFresnelResponseBandLimit(dfield& h, size_t ny, size_t nx, PREC_T dy, PREC_T dx,
                               PREC_T lambda, PREC_T z, PREC_T p_limit) {
    const int64_t lim = detail::win(std::fabs(lambda * z * 0.5 / p_limit) / dx);
    detail::check(olx_fresnel_response_c64(detail::ctx(), h.get(), (int64_t)ny, (int64_t)nx, dy,
                                           dx, lambda, z, lim, lim),
                  "gFresnelResponseBandLimit");
}

// include/golprop.h:506-534
template <typename PREC_T>
void gAsmProp(dfield& u, int ny, int nx, PREC_T dy, PREC_T dx, PREC_T lambda, PREC_T d) {
    detail::check(olx_asm_c64(detail::ctx(), u.get(), ny, nx, dy, dx, lambda, d, 0), "gAsmProp");
}
// include/golprop.h:167-198
template <typename PREC_T>
void gFresnelProp(dfield& u, int ny, int nx, PREC_T dy, PREC_T dx, PREC_T lambda, PREC_T d) {
    detail::check(olx_fresnel_c64(detail::ctx(), u.get(), ny, nx, dy, dx, lambda, d, 0),
                  "gFresnelProp");
}
// include/golprop.h:258-302 (x-limit on both axes)
template <typename PREC_T>
void gFresnelPropBandLimit(dfield& u, int ny, int nx, PREC_T dy, PREC_T dx, PREC_T lambda,
                           PREC_T d) {
    detail::check(olx_fresnel_c64(detail::ctx(), u.get(), ny, nx, dy, dx, lambda, d,
                                  OLX_FRESNEL_BL_GPU), "gFresnelPropBandLimit");
}
// include/golprop.h:578-604
template <typename PREC_T>
void Prop(dfield& u, int ny, int nx, PREC_T dy, PREC_T dx, PREC_T lambda, PREC_T d,
          PROPMODE propmode = PROPMODE::AUTO) {
    const int mode = propmode == ASM ? OLX_MODE_ASM : propmode == FRESNEL ? OLX_MODE_FRESNEL
                                                                          : OLX_MODE_AUTO;
    detail::check(olx_prop_c64(detail::ctx(), u.get(), ny, nx, dy, dx, lambda, d, mode, nullptr),
                  "Prop");
}
// include/golprop.h:941-995; usrc and udst may be the same object; udst is
// allocated when empty
template <typename PREC_T = float>
void ARSSFresnelProp(dfield& usrc, dfield& udst, int ny, int nx, PREC_T dy, PREC_T dx,
                     PREC_T lambda, PREC_T z, PREC_T s, PREC_T oy, PREC_T ox) {
    if (!udst) udst = make_device<std::complex<float>[]>((size_t)ny * nx);
    detail::check(olx_arss_c64(detail::ctx(), usrc.get(), udst.get(), ny, nx, dy, dx, lambda, z,
                               s, oy, ox, 0), "ARSSFresnelProp");
}
template <typename PREC_T = float>
void ARSSFresnelProp(dfield& usrc, dfield& udst, int ny, int nx, PREC_T dy, PREC_T dx,
                     PREC_T lambda, PREC_T z, PREC_T s) {
    ARSSFresnelProp(usrc, udst, ny, nx, dy, dx, lambda, z, s, (PREC_T)0, (PREC_T)0);
}

// include/golprop.h:536-576: shifted ASM without a window, net scale 1; udst is
// allocated when empty and may be usrc
template <typename PREC_T>
void gshiftedAsmProp(dfield& usrc, dfield& udst, int ny, int nx, PREC_T dy, PREC_T dx,
                     PREC_T lambda, PREC_T d, PREC_T oy, PREC_T ox) {
    if (!udst) udst = make_device<std::complex<float>[]>((size_t)ny * nx);
    detail::check(olx_shifted_asm_c64(detail::ctx(), usrc.get(), udst.get(), ny, nx, dy, dx,
                                      lambda, d, oy, ox, 0), "gshiftedAsmProp");
}
// include/golprop.h:756-770: gshiftedAsmProp (ASM) or ARSS with s = 1 (FRESNEL)
template <typename PREC_T>
void shiftedProp(dfield& usrc, dfield& udst, int ny, int nx, PREC_T dy, PREC_T dx,
                 PREC_T lambda, PREC_T d, PREC_T oy, PREC_T ox, PROPMODE propmode) {
    if (propmode == ASM) gshiftedAsmProp(usrc, udst, ny, nx, dy, dx, lambda, d, oy, ox);
    else if (propmode == FRESNEL)
        ARSSFresnelProp(usrc, udst, ny, nx, dy, dx, lambda, d, (PREC_T)1, oy, ox);
    else throw std::runtime_error("shiftedProp: propmode must be ASM or FRESNEL");
}

// the adjoint of AsmProp (conjugate transfer function), for gradient-based design
template <typename PREC_T>
void gAsmPropAdjoint(dfield& u, int ny, int nx, PREC_T dy, PREC_T dx, PREC_T lambda, PREC_T d,
                     bool bandlimit = false) {
    detail::check(olx_asm_c64(detail::ctx(), u.get(), ny, nx, dy, dx, lambda, d,
                              OLX_ADJOINT | (bandlimit ? OLX_BANDLIMIT : 0u)), "gAsmPropAdjoint");
}

}  // namespace ol
#endif
