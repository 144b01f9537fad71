/* Drives every oracle entry point on small inputs (power-of-two and mixed-radix
 * sides, windows active and not, empty clouds, degenerate arguments) so that a build
 * with -fsanitize=address,undefined checks the restatement's memory accesses and
 * arithmetic.  Test infrastructure: built and run by `make -C oracle sanitize`. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "olref.h"

static int failures = 0;
#define EXPECT(cond, what)                                  \
    do {                                                    \
        if (!(cond)) {                                      \
            fprintf(stderr, "FAILED: %s (%s:%d)\n", what, __FILE__, __LINE__); \
            failures++;                                     \
        }                                                   \
    } while (0)

static uint64_t rng_state = 0x9e3779b97f4a7c15ull;
static double urand(void) {
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return (double)(rng_state >> 11) * (1.0 / 9007199254740992.0);
}

static float* field_f(int64_t n) {
    float* u = malloc(sizeof(float) * 2 * (size_t)n);
    for (int64_t i = 0; i < 2 * n; i++) u[i] = (float)(urand() - 0.5);
    return u;
}
static double* field_d(int64_t n) {
    double* u = malloc(sizeof(double) * 2 * (size_t)n);
    for (int64_t i = 0; i < 2 * n; i++) u[i] = urand() - 0.5;
    return u;
}

int main(void) {
    const double lam = 532e-9, p = 8e-6;
    const int64_t shapes[][2] = {{8, 8}, {16, 32}, {6, 10}, {15, 26}, {2, 2}, {7, 11}};
    for (size_t s = 0; s < sizeof shapes / sizeof shapes[0]; s++) {
        const int64_t ny = shapes[s][0], nx = shapes[s][1], n = ny * nx;
        float* f = field_f(n);
        double* d = field_d(n);
        EXPECT(olref_fft2d_c64(f, ny, nx, -1) == 0, "fft c64");
        EXPECT(olref_fft2d_c128(d, ny, nx, +1) == 0, "fft c128");
        EXPECT(olref_asm_c64(f, ny, nx, p, p, lam, 0.01, OLREF_BANDLIMIT) == 0, "asm c64");
        EXPECT(olref_asm_c128(d, ny, nx, p, 6e-6, lam, 0.002, OLREF_ADJOINT) == 0, "asm c128");
        EXPECT(olref_asm_c64(f, ny, nx, p, p, lam, 0.01, OLREF_PREC_F32) == 0, "asm f32");
        EXPECT(olref_fresnel_c64(f, ny, nx, p, p, lam, 0.001, OLREF_FRESNEL_BL_CPU) == 0, "fresnel");
        EXPECT(olref_fresnel_c128(d, ny, nx, p, p, lam, -0.2, OLREF_FRESNEL_BL_GPU) == 0, "fresnel c128");
        float* g = field_f(n);
        EXPECT(olref_arss_c64(f, g, ny, nx, p, p, lam, 0.0006, 0.25, 1e-5, -1e-5, 0) == 0, "arss");
        EXPECT(olref_arss_c64(g, g, ny, nx, p, p, lam, 0.1, 1.0, 0, 0, 0) == 0, "arss alias");
        EXPECT(olref_shifted_fresnel_c64(f, ny, nx, p, p, lam, 0.1, 0.5, 1e-4, 0) == 0, "sfresnel");
        EXPECT(olref_shifted_asm_c64(f, g, ny, nx, p, p, lam, 0.01, 1e-5, 2e-5,
                                     OLREF_SHIFT_WINDOW | OLREF_SHIFT_CPU_SCALE) == 0, "sasm");
        if (ny % 2 == 0 && nx % 2 == 0)
            EXPECT(olref_split_prop_c64(f, ny, nx, p, p, lam, 0.01, 0, 2, 2, 0) == 0, "split");
        EXPECT(olref_tilted_c128(d, ny, nx, p, p, lam, 0.01, 0.1, 1, 0, 0) == 0, "tilted");
        int chosen = -1;
        EXPECT(olref_prop_c64(f, ny, nx, p, p, lam, 0.05, 2, &chosen) == 0 && chosen >= 0, "prop");
        free(f);
        free(g);
        free(d);
    }
    /* point clouds: empty, one point, a few hundred; pixels lists */
    const int64_t np = 300, ny = 24, nx = 20;
    double *x = malloc(8 * np), *y = malloc(8 * np), *z = malloc(8 * np);
    double *a = malloc(8 * np), *ph = malloc(8 * np);
    for (int64_t j = 0; j < np; j++) {
        x[j] = urand() * nx * p;
        y[j] = urand() * ny * p;
        z[j] = 0.151 + 0.004 * urand();
        a[j] = 0.5 + 0.5 * urand();
        ph[j] = 6.28 * urand();
    }
    float* U = malloc(sizeof(float) * 2 * ny * nx);
    double* Ud = malloc(sizeof(double) * 2 * ny * nx);
    for (int64_t cnt = 0; cnt <= np; cnt += np / 2 + 1) {
        EXPECT(olref_cgh_c64(x, y, z, a, ph, cnt, U, ny, nx, p, p, lam, 0) == 0, "cgh");
        EXPECT(olref_cgh_c128(x, y, z, NULL, NULL, cnt, Ud, ny, nx, p, p, lam, OLREF_CGH_EXACT_R) == 0,
               "cgh c128");
        EXPECT(olref_wrp_c64(x, y, z, cnt, U, ny, nx, p, lam, 0.15, p, 0) == 0, "wrp");
        EXPECT(olref_wrp_c128(x, y, z, cnt, Ud, ny, nx, p, lam, 0.15, 2 * p, OLREF_WRP_SQUARE) == 0,
               "wrp square");
        EXPECT(olref_cgh_fresnel_diff(x, y, z, NULL, NULL, cnt, ny, nx, p, lam, Ud) == 0, "diff");
    }
    int64_t hh[5] = {0, 3, 23, 12, 7}, ww[5] = {0, 19, 4, 10, 7};
    double out[10];
    EXPECT(olref_cgh_pixels_c128(x, y, z, a, ph, np, hh, ww, 5, out, p, p, lam, 0) == 0, "pixels");
    EXPECT(olref_wrp_pixels_c128(x, y, z, np, hh, ww, 5, out, p, lam, 0.15, p, 0) == 0, "wrp pixels");
    (void)olref_wrp_plane_z(ny, nx, p, lam, 0.151, p, -1.0f, 0);
    (void)olref_wrp_plane_z(ny, nx, p, lam, 0.151, p, 0.9f, 1);
    (void)olref_asm_prop_check(ny, nx, p, p, lam, 0.1);
    (void)olref_fresnel_prop_check(ny, nx, p, p, lam, 0.1);
    /* optics: diffuser and lens forms */
    float* f = field_f(ny * nx);
    float* b = malloc(sizeof(float) * ny * nx);
    for (int64_t i = 0; i < ny * nx; i++) b[i] = (float)urand();
    EXPECT(olref_diffuser_random_phase(f, ny, nx, 1, 5489u, 6.28f, 1, 2) == 0, "diffuser");
    for (int kind = 0; kind < 6; kind++)
        EXPECT(olref_diffuser_apply(f, ny, nx, b, 1.0f, kind, 4) == 0, "diffuser kinds");
    float* blk = malloc(sizeof(float) * 2 * 16 * 16);
    EXPECT(olref_lens_block(blk, 16, 8e-6f, 0.01f, 532e-9f, 1, 60e-6f) == 0, "lens block");
    for (int kind = 0; kind < 3; kind++)
        EXPECT(olref_lens_tile(blk, 16, f, ny, nx, -3, 2, 8e-6f, 0.01f, 532e-9f, 60e-6f, kind) == 0,
               "lens tile");
    EXPECT(olref_lens_array(f, ny, nx, 8e-6f, 1.28e-4f, 0.01f, 532e-9f, 0, 0, 0, 0.0f, 1) == 0, "lens array");
    (void)olref_mt19937_nth(5489u, 1000);
    double* H = malloc(sizeof(double) * 2 * ny * nx);
    EXPECT(olref_shifted_asm_transfer_c128(H, ny, nx, 1e3, 1e3, lam, 0.01, 1e-5, 0, OLREF_SHIFT_WINDOW) == 0,
           "transfer");
    /* invalid arguments are rejected, not read */
    EXPECT(olref_asm_c64(NULL, 8, 8, p, p, lam, 0.1, 0) != 0, "null field");
    EXPECT(olref_fresnel_c64(f, 8, 8, p, p, lam, 0.0, 0) != 0, "z = 0");
    free(x); free(y); free(z); free(a); free(ph); free(U); free(Ud); free(f); free(b); free(blk); free(H);
    if (failures) {
        fprintf(stderr, "%d failures\n", failures);
        return 1;
    }
    printf("oracle sanitize run: ok\n");
    return 0;
}
