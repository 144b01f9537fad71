// Lensarry.h -- drop-in for the lens arrays of the reference's include/Lensarry.h over
// libolx.so.  The lens field is computed per pixel on the device (the reference
// builds one lenslet with FresnelResponse and tiles it, include/Lensarry.h:137-160);
// host overloads stage the field.  Not provided: duplicate / duplicate_random /
// duplicate_multrandom as public helpers (the random forms draw a std::mt19937
// stream in block-traversal order, :162-252) and the unused shift (:324-348).
#ifndef LENSEARRY_H
#define LENSEARRY_H
#include <cmath>
#include <complex>
#include <cstdio>
#include <memory>

#include "oldefine.h"
#include "oldevice.h"
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

    // include/Lensarry.h:254-265 (host field) and :362-373 (device field)
    virtual void gen_complex(HOLOGRAMSTEP step = PLAY, int oy = 0, int ox = 0) {
        gen(step, oy, ox, 0.f);
    }
    virtual void ggen_complex(HOLOGRAMSTEP step = PLAY, int oy = 0, int ox = 0) { gen(step, oy, ox, 0.f); }
    // include/Lensarry.h:300-311
    void gen_complex_limit(PREC_T p_limit, HOLOGRAMSTEP step = PLAY, int oy = 0, int ox = 0) {
        gen(step, oy, ox, (float)p_limit);
    }
    void ggen_complex_limit(PREC_T p_limit, HOLOGRAMSTEP step = PLAY, int oy = 0, int ox = 0) {
        gen(step, oy, ox, (float)p_limit);
    }
    // include/Lensarry.h:266-271, 374-379: u times the lens field, in place
    virtual void mult_complex(std::unique_ptr<std::complex<PREC_T>[]>& uf, HOLOGRAMSTEP step = PLAY,
                              int oy = 0, int ox = 0) {
        host_mult(uf, step, oy, ox, 0.f);
    }
    virtual void mult_complex(device_ptr<std::complex<float>[]>& uf, HOLOGRAMSTEP step = PLAY,
                              int oy = 0, int ox = 0) {
        run(uf.get(), step, oy, ox, 0.f, 0u);
    }
    // include/Lensarry.h:312-317, 393-398
    void mult_complex_limit(std::unique_ptr<std::complex<PREC_T>[]>& uf, PREC_T p_limit,
                            HOLOGRAMSTEP step = PLAY, int oy = 0, int ox = 0) {
        host_mult(uf, step, oy, ox, (float)p_limit);
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
    void host_mult(std::unique_ptr<std::complex<PREC_T>[]>& uf, HOLOGRAMSTEP st, int oy, int ox,
                   float p_limit) {
        static_assert(std::is_same<PREC_T, float>::value, "Lensarry: complex<float> fields");
        step = st;
        const size_t bytes = sizeof(std::complex<float>) * (size_t)height * width;
        void* d = detail::staging(0).get(bytes);
        detail::check(olx_copy_h2d(detail::ctx(), d, uf.get(), bytes), "Lensarry");
        run(d, st, oy, ox, p_limit, 0u);
        detail::check(olx_copy_d2h(detail::ctx(), uf.get(), d, bytes), "Lensarry");
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
