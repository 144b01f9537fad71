// Diffuser.h -- drop-in for the reference's include/Diffuser.h over libolx.so.
//
// random_phase on a host field is the reference's CPU form: the std::mt19937(seed)
// stream (generated on the host, as there), applied on the device.  The device-field
// methods are the reference's GPU forms; their uniforms come from a Philox4x32-10
// stream (olx_uniform_f32) because cuRAND's MT19937 ordering cannot be reproduced,
// and the shift of random_phase covers the whole field (the reference's shift kernel
// runs on a quarter of it, include/Diffuser.h:151-158).
#ifndef DIFFUSER_H
#define DIFFUSER_H
#include <cmath>
#include <complex>
#include <memory>

#include "oldefine.h"
#include "oldevice.h"
#include "olx.h"

namespace ol {

template <typename _Tp = float>
class Diffuser {
   public:
    using dfield_t = device_ptr<std::complex<float>[]>;
    int pitch = 0;
    int width;
    int height;
    std::unique_ptr<std::complex<_Tp>[]> u;
    Diffuser(int height, int width, int pitch) : pitch(pitch), width(width), height(height) {}
    Diffuser(int height, int width) : width(width), height(height) {}

    // include/Diffuser.h:191-215
    template <typename PREC_T>
    void random_phase(std::unique_ptr<std::complex<PREC_T>[]>& u, HOLOGRAMSTEP step, int seed = 1,
                      float range = 2 * M_PI, int oy = 0, int ox = 0) {
        static_assert(std::is_same<PREC_T, float>::value, "random_phase: complex<float> fields");
        const size_t bytes = sizeof(std::complex<float>) * (size_t)height * width;
        void* d = detail::staging(0).get(bytes);
        detail::check(olx_copy_h2d(detail::ctx(), d, u.get(), bytes), "random_phase");
        detail::check(olx_diffuser_random_phase_c64(detail::ctx(), d, height, width,
                                                    step == PLAY ? OLX_STEP_PLAY : OLX_STEP_RECORD,
                                                    (uint32_t)seed, range, oy, ox), "random_phase");
        detail::check(olx_copy_d2h(detail::ctx(), u.get(), d, bytes), "random_phase");
    }

    // include/Diffuser.h:218-238 (KernelMultPhase over uniforms)
    void random_phase(dfield_t& u, HOLOGRAMSTEP step, int seed = 1, float range = 2 * M_PI) {
        apply(u, OLX_DIFFUSER_PHASE, (size_t)height * width, seed, step == PLAY ? range : -range, 1);
    }
    // include/Diffuser.h:240-259
    void random_xphase(dfield_t& u, HOLOGRAMSTEP step, int seed = 1, float range = 2 * M_PI) {
        apply(u, OLX_DIFFUSER_XPHASE, (size_t)width, seed, step == PLAY ? range : -range, 1);
    }
    // include/Diffuser.h:261-279
    void random_blockphase(dfield_t& u, HOLOGRAMSTEP step, int seed = 1, float range = 2 * M_PI,
                           int blocksize = 1) {
        apply(u, OLX_DIFFUSER_BLOCK, (size_t)width * height / (blocksize * blocksize), seed,
              step == PLAY ? range : -range, blocksize);
    }
    // include/Diffuser.h:281-299
    void random_blockstepphase(dfield_t& u, HOLOGRAMSTEP step, int seed = 1, float range = 2 * M_PI,
                               int blocksize = 4) {
        apply(u, OLX_DIFFUSER_BLOCKSTEP, (size_t)width * height / (blocksize * blocksize), seed,
              step == PLAY ? range : -range, blocksize);
    }
    // include/Diffuser.h:303-314
    void bitphase(dfield_t& u) { apply(u, OLX_DIFFUSER_BITPHASE, 1, 0, 0.f, 1); }
    void phasefunc(dfield_t& u) { bitphase(u); }

   private:
    void apply(dfield_t& u, int kind, size_t nb, int seed, float scalar, int blocksize) {
        auto b = make_device<float[]>(nb);
        detail::check(olx_uniform_f32(detail::ctx(), b.get(), (int64_t)nb, (uint64_t)seed), "uniform");
        detail::check(olx_diffuser_apply_c64(detail::ctx(), u.get(), height, width, b.get(), scalar,
                                             kind, blocksize), "Diffuser");
        synchronize();
    }
};

}  // namespace ol
#endif
