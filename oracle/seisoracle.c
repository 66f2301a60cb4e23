/*
 * seisoracle.c — CPU oracle for the acoustic FD engine.
 *
 * TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg, always as the checker, never as the thing
 * measured or shipped.  The product (seismic_amd/) never links it.
 *
 * It restates the reference (esiwgnahz/seismic, libs/libwave + progs/) in
 * plain C: the same pass order (stencil, injection, 5-level update, Higdon,
 * pointer rotation), the same loop order and the same left-to-right
 * arithmetic.  seisoracle_impl.h is compiled twice:
 *   orc_f_*  REAL=float   (FP32, the reference's precision)
 *   orc_d_*  REAL=double  (FP64 oracle for image/gradient error statements)
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no -ffast-math).
 *
 * Parity pins (tests/test_oracle_*.py):
 *   - kernels and composed forward/backward propagation: bitwise or to a few
 *     ulps against the reference's own Fortran kernels compiled by flang into
 *     oracle/_ref (oracle/ref_harness.f90), composed in Propagator.f90 order;
 *   - FD coefficients: polynomial exactness (degree <= 9), closed form;
 *   - mksinc: shift-0 delta, unit sum; Higdon_coeff: closed formula;
 *   - imaging / FWI gradient / Born: brute-force sums over captured fields;
 *     Born against the central difference of the pinned forward modeling;
 *     LSRTM imaging against the exact discrete Born adjoint; the FWI
 *     gradient against central differences of its misfit (illupow 0 and 1);
 *     the illumination normalisation by its scale invariance and closed form
 *     (tests/test_oracle_pins_imaging.py, tests/test_oracle_gradient.py);
 *   - triangle smoothing: impulse -> triangle, constant preservation;
 *   - L-BFGS (oracle/lbfgs_oracle.c): trajectory equal to lbfgs.f via _ref;
 *   - window/pad helpers (shot_window, window_pad, copy_image), which restate
 *     ExtractPadModel.f90/ReadSouVelRho.f90 (they need SEPlib and cannot be
 *     built here): a table of windows, apertures, off-grid and off-model
 *     columns derived by hand from the geometry.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include "seisfd.h"

typedef struct {
    int nmin1, nmax1, nmin2, nmax2; /* FD region, Fortran indices          */
    int NZ, NX;                     /* allocated incl. 4-cell operator halo */
} ORC_BOUNDS;

/* index of Fortran (i,j) in an array allocated (nmin1-4:nmax1+4, nmin2-4:nmax2+4) */
#define ORC_IDX(b, i, j) ((size_t)((i) - ((b)->nmin1 - 4)) + (size_t)((j) - ((b)->nmin2 - 4)) * (size_t)(b)->NZ)

/* read_window_vel bounds, ExtractPadModel.f90:228-239 (shot/surf type 0) */
void orc_bounds(ORC_BOUNDS *b, int nz, int nxw, int ntaper)
{
    b->nmin1 = -ntaper + 1;
    b->nmax1 = nz + ntaper;
    b->nmin2 = -ntaper + 1;
    b->nmax2 = nxw + ntaper;
    b->NZ = b->nmax1 - b->nmin1 + 9;
    b->NX = b->nmax2 - b->nmin2 + 9;
}

#define REAL float
#define SFX f
#define ROUND roundf
#define FABS fabsf
#define SIN sinf
#define COS cosf
#define POW powf
#define SQRT sqrtf
#define LOG logf
#include "seisoracle_impl.h"
#undef REAL
#undef SFX
#undef ROUND
#undef FABS
#undef SIN
#undef COS
#undef POW
#undef SQRT
#undef LOG

#define REAL double
#define SFX d
#define ROUND round
#define FABS fabs
#define SIN sin
#define COS cos
#define POW pow
#define SQRT sqrt
#define LOG log
#include "seisoracle_impl.h"

/* Mute.f90:50-98 MuteParam_compute_mask with compute_top_mute (:171-195) and
 * compute_bottom_mute (:197-219), literally: per trace mute = 0, the sin^2
 * loop, mute(it+1:) = 1; then 1, the cos^2 loop, mute(it+1+nramp:) = 0; the
 * mask is their product.  hy = 0 (2D).  mask: per shot (nt, ntr) blocks. */
static void mute_top(float v0, float t0m, float tramp, float dt, float ot, float hx, int nt, float *m)
{
    const float pi = 3.1415926535f;
    int nramp = (int)(tramp / dt + 1.5f);
    float q = hx / v0;
    float tm = sqrtf(t0m * t0m + q * q);
    int it = (int)((tm - ot) / dt + 1.5f);
    for (int i = 1; i <= nt; i++) m[i - 1] = 0.f;
    int beg = it - nramp > 1 ? it - nramp : 1, end = it < nt ? it : nt;
    for (int i = beg; i <= end; i++) {
        float s = sinf((float)(i - it + nramp) * pi / (float)(2 * nramp));
        m[i - 1] = s * s;
    }
    for (int i = (it + 1 > 1 ? it + 1 : 1); i <= nt; i++) m[i - 1] = 1.f;
}

static void mute_bottom(float v1, float t1m, float tramp, float dt, float ot, float hx, int nt, float *m)
{
    const float pi = 3.1415926535f;
    int nramp = (int)(tramp / dt + 1.5f);
    float q = hx / v1;
    float tm = sqrtf(t1m * t1m + q * q);
    int it = (int)((tm - ot) / dt + 1.5f);
    for (int i = 1; i <= nt; i++) m[i - 1] = 1.f;
    int beg = it + 1 > 1 ? it + 1 : 1, end = it + 1 + nramp < nt ? it + 1 + nramp : nt;
    for (int i = beg; i <= end; i++) {
        float c = cosf((float)(i - it - 1) * pi / (float)(2 * nramp));
        m[i - 1] = c * c;
    }
    for (int i = (it + 1 + nramp > 1 ? it + 1 + nramp : 1); i <= nt; i++) m[i - 1] = 0.f;
}

void orc_f_mute_mask(int nshots, const sfd_shot *shots, float dt, float ot, const float *opt5, float *mask)
{
    /* opt5: v0, t0, v1, t1, tramp (sfd_mute_opts order) */
    size_t off = 0;
    for (int s = 0; s < nshots; s++) {
        const sfd_shot *sh = &shots[s];
        float *a = (float *)malloc(sizeof(float) * sh->nt), *b = (float *)malloc(sizeof(float) * sh->nt);
        for (int r = 0; r < sh->ntr; r++) {
            float hx = sh->rx[r] - sh->sx;
            mute_top(opt5[0], opt5[1], opt5[4], dt, ot, hx, sh->nt, a);
            mute_bottom(opt5[2], opt5[3], opt5[4], dt, ot, hx, sh->nt, b);
            for (int i = 0; i < sh->nt; i++) mask[off + i] = a[i] * b[i];
            off += (size_t)sh->nt;
        }
        free(a);
        free(b);
    }
}
