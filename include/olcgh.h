// olcgh.h -- drop-in for the reference's CGH / WRP API on MI355X.
//
// Host overloads keep the signatures of include/olcgh.h (ol::Cgh, ol::WRPStep1);
// device overloads mirror include/golcgh.h (gCgh, WRPStep1(_D), WRPStep2,
// WRPMethod(_D)) on ol::gObject and ol::device_ptr.  All sums run in the HIP
// kernels of libolx.so; errors throw std::runtime_error.
#ifndef OLCGH_H
#define OLCGH_H
#include <cmath>
#include <complex>
#include <memory>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "olclass.h"
#include "oldefine.h"
#include "oldevice.h"
#include "olprop.h"

namespace ol {
namespace detail {

template <typename _Tp>
void* upload_points(const Object<_Tp>& obj) {
    auto v = to_points(obj);
    void* d = staging(1).get(sizeof(olx_point) * (v.empty() ? 1 : v.size()));
    if (!v.empty()) check(olx_copy_h2d(ctx(), d, v.data(), sizeof(olx_point) * v.size()), "points");
    return d;
}

// complex hologram on the device for obj, copied into u (complex<C>)
template <typename _Tp, typename C>
void cgh_complex(const Object<_Tp>& obj, std::unique_ptr<std::complex<C>[]>& u, int ny, int nx,
                 double py, double px, double lambda, uint32_t flags = 0) {
    void* pts = upload_points(obj);
    on_device(u, (int64_t)ny * nx, [&](void* d) {
        check(olx_cgh_c64(ctx(), (const olx_point*)pts, obj.size, d, nullptr, nullptr, ny, nx, py,
                          px, lambda, flags), "Cgh");
    });
}

}  // namespace detail

// phase-only output (atan2 of the sum), include/olcgh.h:351-365.
template <typename _Tp, typename PHASE_T, typename PREC_T>
void Cgh(Object<_Tp>& obj, std::unique_ptr<PHASE_T[]>& phase, int ny, int nx, PREC_T py,
         PREC_T px, PREC_T lambda) {
    static_assert(std::is_floating_point<PHASE_T>::value, "phase output must be float or double");
    void* pts = detail::upload_points(obj);
    const size_t n = (size_t)ny * nx;
    void* d = detail::staging(0).get(sizeof(float) * n);
    detail::check(olx_cgh_c64(detail::ctx(), (const olx_point*)pts, obj.size, nullptr, (float*)d,
                              nullptr, ny, nx, py, px, lambda, detail::prec_flags<PREC_T>()), "Cgh");
    if constexpr (std::is_same<PHASE_T, float>::value) {
        detail::check(olx_copy_d2h(detail::ctx(), phase.get(), d, sizeof(float) * n), "Cgh");
    } else {
        std::vector<float> t(n);
        detail::check(olx_copy_d2h(detail::ctx(), t.data(), d, sizeof(float) * n), "Cgh");
        for (size_t i = 0; i < n; i++) phase[i] = (PHASE_T)t[i];
    }
}
// normalized = true runs _CghFresnelNormalized (include/olcgh.h:213-231): point
// coordinates in pixels, dx = p (w - x_j), dy = p (h - x_j) (the reference reads x_j
// for both axes), dz = p z_j, unit amplitude.  That is the plain Fresnel CGH of the
// cloud (p x_j, p x_j, p z_j), which is what runs here.
template <typename _Tp, typename PHASE_T, typename PREC_T>
void Cgh(Object<_Tp>& obj, std::unique_ptr<PHASE_T[]>& phase, int ny, int nx, PREC_T p,
         PREC_T lambda, bool normalized = false) {
    if (!normalized) {
        Cgh(obj, phase, ny, nx, p, p, lambda);
        return;
    }
    const double pd = (double)p;
    Object<double> scaled(obj.size);
    for (int j = 0; j < obj.size; j++)
        scaled[j] = Vect3<double>(pd * (double)obj[j].x, pd * (double)obj[j].x, pd * (double)obj[j].z);
    Cgh(scaled, phase, ny, nx, p, p, lambda);
}

// include/olcgh.h:315-349: the recurrence-formula CGH, coordinates in pixel units
// (theta = p z / lambda + gamma / 2 (dx^2 + dy^2) cycles, gamma = p / (lambda z)),
// phase output.  The reference's theta and delta are uninitialised for w > 0; the
// device evaluates the recurrence's exact value at every pixel.
template <typename _Tp, typename PHASE_T, typename PREC_T>
void _CghFresnelDiff(Object<_Tp>& obj, std::unique_ptr<PHASE_T[]>& phase, int ny, int nx, PREC_T p,
                     PREC_T lambda) {
    static_assert(std::is_floating_point<PHASE_T>::value, "phase output must be float or double");
    void* pts = detail::upload_points(obj);
    const size_t n = (size_t)ny * nx;
    void* d = detail::staging(0).get(sizeof(float) * n);
    detail::check(olx_cgh_fresnel_diff_c64(detail::ctx(), (const olx_point*)pts, obj.size, nullptr,
                                           (float*)d, ny, nx, p, lambda), "_CghFresnelDiff");
    if constexpr (std::is_same<PHASE_T, float>::value) {
        detail::check(olx_copy_d2h(detail::ctx(), phase.get(), d, sizeof(float) * n), "_CghFresnelDiff");
    } else {
        std::vector<float> t(n);
        detail::check(olx_copy_d2h(detail::ctx(), t.data(), d, sizeof(float) * n), "_CghFresnelDiff");
        for (size_t i = 0; i < n; i++) phase[i] = (PHASE_T)t[i];
    }
}

// complex output, include/olcgh.h:367-379 (_CghFresnel :296-313)
template <typename _Tp, typename COMPLEX_T, typename PREC_T>
void Cgh(Object<_Tp>& obj, std::unique_ptr<std::complex<COMPLEX_T>[]>& u, int ny, int nx, PREC_T p,
         PREC_T lambda) {
    detail::cgh_complex(obj, u, ny, nx, p, p, lambda, detail::prec_flags<PREC_T>());
}
template <typename _Tp, typename COMPLEX_T, typename PREC_T>
void Cgh(Object<_Tp>& obj, std::unique_ptr<std::complex<COMPLEX_T>[]>& u, int ny, int nx,
         PREC_T py, PREC_T px, PREC_T lambda) {
    detail::cgh_complex(obj, u, ny, nx, py, px, lambda, detail::prec_flags<PREC_T>());
}

// include/olcgh.h:381-399: the CPU WRPStep1 has no window and ignores d (its sum is
// the plain Fresnel CGH); kept as the reference defines it
template <typename _Tp, typename COMPLEX_T, typename PREC_T>
void WRPStep1(Object<_Tp>& obj, std::unique_ptr<std::complex<COMPLEX_T>[]>& u, int ny, int nx,
              PREC_T p, PREC_T lambda, PREC_T d) {
    (void)d;
    detail::cgh_complex(obj, u, ny, nx, p, p, lambda);
}

// ------------------------------------------------------------------ device API
// include/golcgh.h:127-145
template <typename PREC_T>
void gCgh(gObject<PREC_T>& gobj, dfield& U, int ny, int nx, PREC_T py, PREC_T px, PREC_T lambda) {
    detail::check(olx_cgh_c64(detail::ctx(), gobj.points.get(), gobj.size, U.get(), nullptr, nullptr,
                              ny, nx, py, px, lambda, 0), "gCgh");
}
template <typename PREC_T>
void gCgh(gObject<PREC_T>& gobj, dfield& U, int ny, int nx, PREC_T p, PREC_T lambda) {
    gCgh(gobj, U, ny, nx, p, p, lambda);
}
// phase-only device output (gCgh with a real CGH_T buffer, include/golcgh.h:5-45)
template <typename PREC_T>
void gCgh(gObject<PREC_T>& gobj, device_ptr<float[]>& phase, int ny, int nx, PREC_T py, PREC_T px,
          PREC_T lambda) {
    detail::check(olx_cgh_c64(detail::ctx(), gobj.points.get(), gobj.size, nullptr, phase.get(),
                              nullptr, ny, nx, py, px, lambda, 0), "gCgh");
}
// extension: exact spherical distance and per-point amplitude / phase (from Object2dev)
template <typename PREC_T>
void gCghExact(gObject<PREC_T>& gobj, dfield& U, int ny, int nx, PREC_T py, PREC_T px,
               PREC_T lambda) {
    detail::check(olx_cgh_c64(detail::ctx(), gobj.points.get(), gobj.size, U.get(), nullptr, nullptr,
                              ny, nx, py, px, lambda, OLX_CGH_EXACT_R), "gCghExact");
}

// include/golcgh.h:369-389 (circle) and :471-491 (square, _D)
template <typename _Tp, typename PREC_T>
void WRPStep1(gObject<_Tp>& obj, dfield& U, int ny, int nx, PREC_T p, PREC_T lambda, PREC_T d) {
    detail::check(olx_wrp_c64(detail::ctx(), obj.points.get(), obj.size, U.get(), ny, nx, p, lambda,
                              d, p, 0), "WRPStep1");
}
template <typename _Tp, typename PREC_T>
void WRPStep1(gObject<_Tp>& obj, dfield& U, int ny, int nx, PREC_T p, PREC_T lambda, PREC_T d,
              PREC_T plimit) {
    detail::check(olx_wrp_c64(detail::ctx(), obj.points.get(), obj.size, U.get(), ny, nx, p, lambda,
                              d, plimit, 0), "WRPStep1");
}
template <typename _Tp, typename PREC_T>
void WRPStep1_D(gObject<_Tp>& obj, dfield& U, int ny, int nx, PREC_T p, PREC_T lambda, PREC_T d) {
    detail::check(olx_wrp_c64(detail::ctx(), obj.points.get(), obj.size, U.get(), ny, nx, p, lambda,
                              d, p, OLX_WRP_SQUARE), "WRPStep1_D");
}
template <typename _Tp, typename PREC_T>
void WRPStep1_D(gObject<_Tp>& obj, dfield& U, int ny, int nx, PREC_T p, PREC_T lambda, PREC_T d,
                PREC_T plimit) {
    detail::check(olx_wrp_c64(detail::ctx(), obj.points.get(), obj.size, U.get(), ny, nx, p, lambda,
                              d, plimit, OLX_WRP_SQUARE), "WRPStep1_D");
}

// include/golcgh.h:494-517: AUTO uses ASM when AsmPropCheck holds, else Fresnel when
// FresnelPropCheck holds, else ASM
template <typename PREC_T>
void WRPStep2(dfield& U, int ny, int nx, PREC_T p, PREC_T lambda, PREC_T d,
              PROPMODE propmode = PROPMODE::AUTO) {
    if (propmode == AUTO)
        propmode = (AsmPropCheck<double>(ny, nx, p, p, lambda, d) ||
                    !FresnelPropCheck<double>(ny, nx, p, p, lambda, d)) ? ASM : FRESNEL;
    if (propmode == ASM) gAsmProp(U, ny, nx, p, p, lambda, d);
    else gFresnelProp(U, ny, nx, p, p, lambda, d);
}

namespace detail {
template <typename _Tp, typename PREC_T>
double wrp_method(gObject<_Tp>& obj, dfield& U, int ny, int nx, PREC_T p, PREC_T lambda,
                  PREC_T zmin, PREC_T plimit, float s, PROPMODE propmode, uint32_t flags) {
    const int mode = propmode == ASM ? OLX_MODE_ASM : propmode == FRESNEL ? OLX_MODE_FRESNEL
                                                                          : OLX_MODE_AUTO;
    double pz = 0;
    check(olx_wrp_method_c64(ctx(), obj.points.get(), obj.size, U.get(), ny, nx, p, lambda, zmin,
                             plimit, s, mode, flags, &pz), "WRPMethod");
    return pz;
}
}  // namespace detail

// include/golcgh.h:561-581 (circle windows); returns the WRP plane position
template <typename _Tp, typename PREC_T>
double WRPMethod(gObject<_Tp>& obj, dfield& U, int ny, int nx, PREC_T p, PREC_T lambda, PREC_T zmin,
                 float s = 0.9, PROPMODE propmode = PROPMODE::AUTO) {
    return detail::wrp_method(obj, U, ny, nx, p, lambda, zmin, p, s, propmode, 0u);
}
// include/golcgh.h:519-559 (square windows; optional p_limit)
template <typename _Tp, typename PREC_T>
double WRPMethod_D(gObject<_Tp>& obj, dfield& U, int ny, int nx, PREC_T p, PREC_T lambda,
                   PREC_T zmin, float s = 0.9, PROPMODE propmode = PROPMODE::AUTO) {
    return detail::wrp_method(obj, U, ny, nx, p, lambda, zmin, p, s, propmode, OLX_WRP_SQUARE);
}
template <typename _Tp, typename PREC_T>
double WRPMethod_D(gObject<_Tp>& obj, dfield& U, int ny, int nx, PREC_T p, PREC_T lambda,
                   PREC_T zmin, PREC_T p_limit, float s = 0.9, PROPMODE propmode = PROPMODE::AUTO) {
    return detail::wrp_method(obj, U, ny, nx, p, lambda, zmin, p_limit, s, propmode, OLX_WRP_SQUARE);
}

// extension: layered Fresnel CGH on MFMA (precision OLX_LAYERED_F16 / _BF16 / _BF16X3)
template <typename _Tp, typename PREC_T>
void CghLayered(gObject<_Tp>& obj, dfield& U, int ny, int nx, PREC_T py, PREC_T px, PREC_T lambda,
                int nlayers, PREC_T zmin, PREC_T zmax, int precision = OLX_LAYERED_F16) {
    detail::check(olx_cgh_layered_c64(detail::ctx(), obj.points.get(), obj.size, U.get(), ny, nx,
                                      py, px, lambda, nlayers, zmin, zmax, precision, 0),
                  "CghLayered");
}

}  // namespace ol
#endif
