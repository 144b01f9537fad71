// olx_layered.hip -- layered Fresnel CGH as complex GEMMs on MFMA (placeholder until the
// MFMA kernel lands: the entry point reports EUNSUPPORTED).
#include "olx_common.hpp"

using namespace olx;

extern "C" double olx_layer_depth(double z, double zmin, double zmax, int nlayers) {
    const double dz = (zmax - zmin) / nlayers;
    long long l = (long long)std::floor((z - zmin) / (zmax - zmin) * nlayers);
    if (l < 0) l = 0;
    if (l >= nlayers) l = nlayers - 1;
    return zmin + ((double)l + 0.5) * dz;
}

extern "C" int olx_cgh_layered_c64(olx_ctx* ctx, const olx_point* d_pts, int64_t n, void* d_U,
                                   int64_t ny, int64_t nx, double py, double px, double lambda,
                                   int nlayers, double zmin, double zmax, int precision,
                                   uint32_t flags) {
    OLX_FAIL(OLX_EUNSUPPORTED, "layered CGH not built yet");
}
