// Lensarry.h -- drop-in for the lens arrays of the reference's include/Lensarry.h over
// libolx.so.  The lens field is computed per pixel on the device (the reference
// builds one lenslet with FresnelResponse and tiles it, include/Lensarry.h:252-263);
// the random forms tile the lenslet with phases drawn from std::mt19937(0) in the
// reference's traversal order (the host replays the stream, the device tiles).  Host
// mult_complex* keep the reference's structure: the lens field is generated when
// absent or when the step changed, then multiplied in with ol::mul_complex
// (include/Lensarry.h:264-269).  Not provided: the unused shift (:332-360).
#ifndef LENSEARRY_H
#define LENSEARRY_H
#include <cmath>
#include <complex>
#include <cstdio>
#include <memory>

#include "oldefine.h"
#include "oldevice.h"
#include "olutils.h"
#include "olx.h"

namespace ol {

template <typename PREC_T = float>
class Lensarry {
   public:
    float pitch;
    int width;
    int height;
    float f;
    float lambda;
    float blockwidth;
    int blockpx;
    std::unique_ptr<std::complex<PREC_T>[]> u;
    device_ptr<std::complex<float>[]> du;
    HOLOGRAMSTEP step = PLAY;
    // include/Lensarry.h:98-106
    Lensarry(int height, int width, float pitch, float blockwidth, float f, float lambda)
        : pitch(pitch), width(width), height(height), f(f), lambda(lambda), blockwidth(blockwidth),
          blockpx((int)std::round(blockwidth / pitch)) {}
    virtual ~Lensarry() {}
    void info() {
        printf("pitch[um] = %lf\n", pitch * 1e6);
        printf("blockpx = %d\n", blockpx);
    }

    // include/Lensarry.h:252-263 (host field) and :369-380 (device field)
    virtual void gen_complex(HOLOGRAMSTEP step = PLAY, int oy = 0, int ox = 0) {
        gen(step, oy, ox, 0.f);
    }
    virtual void ggen_complex(HOLOGRAMSTEP step = PLAY, int oy = 0, int ox = 0) { gen(step, oy, ox, 0.f); }
    // include/Lensarry.h:290-302, 389-400
    void gen_complex_limit(PREC_T p_limit, HOLOGRAMSTEP step = PLAY, int oy = 0, int ox = 0) {
        gen(step, oy, ox, (float)p_limit);
    }
    void ggen_complex_limit(PREC_T p_limit, HOLOGRAMSTEP step = PLAY, int oy = 0, int ox = 0) {
        gen(step, oy, ox, (float)p_limit);
    }
    // include/Lensarry.h:264-269 (host field): generate when absent or the step changed,
    // then u times the lens field with ol::mul_complex; :381-387 (device field)
    virtual void mult_complex(std::unique_ptr<std::complex<PREC_T>[]>& uf, HOLOGRAMSTEP step = PLAY,
                              int oy = 0, int ox = 0) {
        if (!this->u || this->step != step) gen_complex(step, oy, ox);
        mul_complex(uf, this->u, uf, this->height, this->width);
    }
    virtual void mult_complex(device_ptr<std::complex<float>[]>& uf, HOLOGRAMSTEP step = PLAY,
                              int oy = 0, int ox = 0) {
        run(uf.get(), step, oy, ox, 0.f, 0u);
    }
    // include/Lensarry.h:303-309, 401-407
    void mult_complex_limit(std::unique_ptr<std::complex<PREC_T>[]>& uf, PREC_T p_limit,
                            HOLOGRAMSTEP step = PLAY, int oy = 0, int ox = 0) {
        if (!this->u || this->step != step) gen_complex_limit(p_limit, step, oy, ox);
        mul_complex(uf, this->u, uf, this->height, this->width);
    }
    // include/Lensarry.h:136-161: ublock (blockpx x blockpx) tiled over the field
    std::unique_ptr<std::complex<PREC_T>[]> duplicate(std::unique_ptr<std::complex<PREC_T>[]>& ublock,
                                                      int on = 0, int om = 0) {
        return host_tile(ublock, on, om, 0.f, 0);
    }
    // include/Lensarry.h:163-206: outside the p_limit window exp(i 2 pi (v - 0.5))
    std::unique_ptr<std::complex<PREC_T>[]> duplicate_random(
        std::unique_ptr<std::complex<PREC_T>[]>& ublock, PREC_T p_limit, int on = 0, int om = 0) {
        return host_tile(ublock, on, om, (float)p_limit, 1);
    }
    // include/Lensarry.h:207-250: outside the window the block value times exp(i pi/2 (v - 0.5))
    std::unique_ptr<std::complex<PREC_T>[]> duplicate_multrandom(
        std::unique_ptr<std::complex<PREC_T>[]>& ublock, PREC_T p_limit, int on = 0, int om = 0) {
        return host_tile(ublock, on, om, (float)p_limit, 2);
    }
    // include/Lensarry.h:270-281: the unwindowed lenslet tiled by duplicate_multrandom
    void gen_complex_random(PREC_T p_limit, HOLOGRAMSTEP step = PLAY, int oy = 0, int ox = 0) {
        gen_tile(step, oy, ox, 0.f, (float)p_limit, 2);
    }
    // include/Lensarry.h:282-288
    void mult_complex_random(std::unique_ptr<std::complex<PREC_T>[]>& uf, PREC_T p_limit,
                             HOLOGRAMSTEP step = PLAY, int oy = 0, int ox = 0) {
        if (!this->u || this->step != step) gen_complex_random(p_limit, step, oy, ox);
        mul_complex(uf, this->u, uf, this->height, this->width);
    }
    // include/Lensarry.h:310-321: the band-limited lenslet tiled by duplicate_random
    void gen_complex_limit_random(PREC_T p_limit, HOLOGRAMSTEP step = PLAY, int oy = 0, int ox = 0) {
        gen_tile(step, oy, ox, (float)p_limit, (float)p_limit, 1);
    }
    // include/Lensarry.h:322-328
    void mult_complex_limit_random(std::unique_ptr<std::complex<PREC_T>[]>& uf, PREC_T p_limit,
                                   HOLOGRAMSTEP step = PLAY, int oy = 0, int ox = 0) {
        if (!this->u || this->step != step) gen_complex_limit_random(p_limit, step, oy, ox);
        mul_complex(uf, this->u, uf, this->height, this->width);
    }
    void mult_complex_limit(device_ptr<std::complex<float>[]>& uf, PREC_T p_limit,
                            HOLOGRAMSTEP step = PLAY, int oy = 0, int ox = 0) {
        run(uf.get(), step, oy, ox, (float)p_limit, 0u);
    }

   protected:
    virtual uint32_t shape_flags() const { return 0u; }
    void run(void* d, HOLOGRAMSTEP st, int oy, int ox, float p_limit, uint32_t extra) {
        detail::check(olx_lens_array_c64(detail::ctx(), d, height, width, pitch, blockwidth, f, lambda,
                                         st == PLAY ? OLX_STEP_PLAY : OLX_STEP_RECORD, oy, ox, p_limit,
                                         shape_flags() | extra), "Lensarry");
    }
    // the lens field into du (device) and u (host copy)
    void gen(HOLOGRAMSTEP st, int oy, int ox, float p_limit) {
        step = st;
        const size_t n = (size_t)height * width;
        du = make_device<std::complex<float>[]>(n);
        run(du.get(), st, oy, ox, p_limit, OLX_LENS_WRITE);
        static_assert(std::is_same<PREC_T, float>::value, "Lensarry: complex<float> fields");
        u = dev2cpu(du, n);
    }
    // the block tiled on the device (kind 0 / 1 / 2: duplicate / _random / _multrandom)
    std::unique_ptr<std::complex<PREC_T>[]> host_tile(std::unique_ptr<std::complex<PREC_T>[]>& ublock,
                                                      int on, int om, float p_limit, int kind) {
        static_assert(std::is_same<PREC_T, float>::value, "Lensarry: complex<float> fields");
        auto db = cpu2dev(ublock, (size_t)blockpx * blockpx);
        const size_t n = (size_t)height * width;
        auto dout = make_device<std::complex<float>[]>(n);
        detail::check(olx_lens_tile_c64(detail::ctx(), db.get(), blockpx, dout.get(), height, width, on,
                                        om, pitch, f, lambda, p_limit, kind), "Lensarry duplicate");
        return dev2cpu(dout, n);
    }
    // the lenslet (band limit block_limit > 0) tiled with the random form kind into du and u
    void gen_tile(HOLOGRAMSTEP st, int oy, int ox, float block_limit, float p_limit, int kind) {
        static_assert(std::is_same<PREC_T, float>::value, "Lensarry: complex<float> fields");
        step = st;
        auto block = make_device<std::complex<float>[]>((size_t)blockpx * blockpx);
        detail::check(olx_lens_array_c64(detail::ctx(), block.get(), blockpx, blockpx, pitch, blockwidth,
                                         f, lambda, st == PLAY ? OLX_STEP_PLAY : OLX_STEP_RECORD, 0, 0,
                                         block_limit, OLX_LENS_WRITE), "Lensarry");
        const size_t n = (size_t)height * width;
        du = make_device<std::complex<float>[]>(n);
        detail::check(olx_lens_tile_c64(detail::ctx(), block.get(), blockpx, du.get(), height, width, oy,
                                        ox, pitch, f, lambda, p_limit, kind), "Lensarry");
        u = dev2cpu(du, n);
    }
};

// include/Lensarry.h:383-447: hexagonal rows of circular lenslets on a field of ones
template <typename PREC_T>
class LensarryCircle : public Lensarry<PREC_T> {
   public:
    using Lensarry<PREC_T>::Lensarry;

   protected:
    uint32_t shape_flags() const override { return OLX_LENS_CIRCLE; }
};

}  // namespace ol
#endif
