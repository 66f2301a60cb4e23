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

/* read_window_vel bounds, ExtractPadModel.f90:228-239: a ghost source keeps
 * ntaper + lsinc/2 + 2 FD rows above the model, a free surface lsinc - 1
 * (the surface wins when both are set). */
void orc_bounds_t(ORC_BOUNDS *b, int nz, int nxw, int ntaper, int lsinc, int shot_type, int surf_type)
{
    if (shot_type > 0 || surf_type > 0) {
        if (shot_type > 0) b->nmin1 = -(ntaper + lsinc / 2 + 2) + 1;
        if (surf_type > 0) b->nmin1 = -lsinc + 1;
    } else {
        b->nmin1 = -ntaper + 1;
    }
    b->nmax1 = nz + ntaper;
    b->nmin2 = -ntaper + 1;
    b->nmax2 = nxw + ntaper;
    b->NZ = b->nmax1 - b->nmin1 + 9;
    b->NX = b->nmax2 - b->nmin2 + 9;
}

void orc_bounds(ORC_BOUNDS *b, int nz, int nxw, int ntaper)
{
    orc_bounds_t(b, nz, nxw, ntaper, 7, 0, 0);
}

/* Density per model cell (nz*nx, column-major; the reference's mod%rho with
 * withRho): a copy held for the oracle's calls; NULL or n = 0 restores
 * constant density (withRho false). */
static float *orc_rho = NULL;
static size_t orc_rho_n = 0;
void orc_set_density(const float *r, size_t n)
{
    free(orc_rho);
    orc_rho = NULL;
    orc_rho_n = 0;
    if (!r || n < 1) return;
    orc_rho = (float *)malloc(sizeof(float) * n);
    memcpy(orc_rho, r, sizeof(float) * n);
    orc_rho_n = n;
}

/* Free-surface elevation per model column (ModelSpace_elevation%elev, which
 * the reference's drivers set to 0): a copy held for the oracle's calls;
 * NULL or n = 0 restores 0 everywhere. */
static float *orc_elev = NULL;
static int orc_elev_n = 0;
void orc_set_elevation(const float *e, int n)
{
    free(orc_elev);
    orc_elev = NULL;
    orc_elev_n = 0;
    if (!e || n < 1) return;
    orc_elev = (float *)malloc(sizeof(float) * (size_t)n);
    memcpy(orc_elev, e, sizeof(float) * (size_t)n);
    orc_elev_n = n;
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

/* ---- Butterworth band-pass (libs/libwave/bandpass.c:72-266, called by
 * Bandpass.f90:89-147).  The filters are cascades of first- and second-order
 * sections applied sample by sample in FP32 with coefficients built in the
 * reference's mix of float and double; phase 0 runs every section forward
 * and then over the time-reversed record.  Restated here from the
 * published recipe (Oppenheim and Schafer 1975; Hale's Butterworth
 * sections) and pinned bit for bit to the reference's own bandpass.c,
 * compiled from its source into oracle/_ref. */
static const double bp_pi = 3.14159265358979323846;

static int bp_sections(int *np, int phase)
{
    int nodd = *np % 2;
    if (phase == 0) {
        *np = nodd ? (*np + 1) / 2 : *np / 2;
    }
    return (*np + 1) / 2;
}

/* one pass of every section over x[2..n-1] (y[0], y[1] are never written) */
static void bp_run(const float *c, int nsec, int n, float *x, float *y)
{
    for (int s = 0; s < nsec; s++) {
        const float *b = c + 5 * s;
        for (int i = 2; i < n; i++)
            y[i] = b[0] * x[i] + b[1] * x[i - 1] + b[2] * x[i - 2] - b[3] * y[i - 1] - b[4] * y[i - 2];
        for (int i = 0; i < n; i++) x[i] = y[i];
    }
}

static void bp_apply(const float *c, int nsec, int phase, int n, float *data, float *nw, float *tmp)
{
    bp_run(c, nsec, n, data, nw);
    if (phase == 0) {
        for (int i = 2; i < n; i++) tmp[i] = data[n + 1 - i];
        bp_run(c, nsec, n, tmp, nw);
        for (int i = 2; i < n; i++) data[i] = tmp[n + 1 - i];
    }
}

/* highcut_ (bandpass.c:72-140) coefficients: c[5*s + k], s < nsec */
int orc_highcut_coefs(float fhi, int nphi, int phase, float *c)
{
    int np = nphi, nsec = bp_sections(&np, phase), nodd = np % 2;
    float a = 2.0 * tan(bp_pi * fhi), aa = a * a, aap4 = aa + 4.0;
    float dtheta = bp_pi / np, theta0 = nodd ? 0.0f : dtheta / 2.0;
    if (nodd) {
        c[0] = a / (a + 2.0);
        c[1] = c[0];
        c[2] = 0.0f;
        c[3] = (a - 2.0) / (a + 2.0);
        c[4] = 0.0f;
    }
    for (int j = nodd; j < nsec; j++) {
        double cc = 4.0 * a * cos(theta0 + j * dtheta);
        float *b = c + 5 * j;
        b[0] = aa / (aap4 + cc);
        b[1] = 2.0 * b[0];
        b[2] = b[0];
        b[3] = (2.0 * aa - 8.0) / (aap4 + cc);
        b[4] = (aap4 - cc) / (aap4 + cc);
    }
    return nsec;
}

/* lowcut_ (bandpass.c:185-266): the highcut at half Nyquist moved to flo */
int orc_lowcut_coefs(float flo, int nplo, int phase, float *c)
{
    int np = nplo, nsec = bp_sections(&np, phase), nodd = np % 2;
    float fno2 = 0.25f;
    float a = 2. * sin(bp_pi * fno2) / cos(bp_pi * fno2), aa = a * a, aap4 = aa + 4;
    float e = -cos(bp_pi * (flo + fno2)) / cos(bp_pi * (flo - fno2)), ee = e * e;
    float dtheta = bp_pi / np, theta0 = nodd ? 0 : dtheta / 2;
    if (nodd) {
        float b1 = a / (a + 2), b2 = (a - 2) / (a + 2), den = 1.0 - b2 * e;
        c[0] = b1 * (1. - e) / den;
        c[1] = -c[0];
        c[2] = 0.f;
        c[3] = (e - b2) / den;
        c[4] = 0.f;
    }
    for (int j = nodd; j < nsec; j++) {
        double cc = 4. * a * cos(theta0 + j * dtheta);
        float b1 = aa / (aap4 + cc), b2 = (2. * aa - 8.) / (aap4 + cc), b3 = (aap4 - cc) / (aap4 + cc);
        float den = 1. - b2 * e + b3 * ee;
        float *d = c + 5 * j;
        d[0] = b1 * (1. - e) * (1. - e) / den;
        d[1] = -2. * d[0];
        d[2] = d[0];
        d[3] = (2. * e * (1. + b3) - b2 * (1. + ee)) / den;
        d[4] = (ee - b2 * e + b3) / den;
    }
    return nsec;
}

/* highcut_ / lowcut_ on one array (the reference's calling convention) */
void orc_highcut(float fhi, int nphi, int phase, int n, float *data, float *nw, float *tmp)
{
    float c[5 * 64];
    if (nphi > 128) return;
    int nsec = orc_highcut_coefs(fhi, nphi, phase, c);
    bp_apply(c, nsec, phase, n, data, nw, tmp);
}

void orc_lowcut(float flo, int nplo, int phase, int n, float *data, float *nw, float *tmp)
{
    float c[5 * 64];
    if (nplo > 128) return;
    int nsec = orc_lowcut_coefs(flo, nplo, phase, c);
    bp_apply(c, nsec, phase, n, data, nw, tmp);
}

/* Bandpass.f90:89-147 bandpass: each trace (nt samples, traces nt apart) is
 * padded by npad = 10 samples each side with sin^2 tapers of its end values,
 * low-cut when flo dt > 1e-6, high-cut when fhi dt < 4.999999, and the nt
 * samples after the front pad are kept.  Returns -1 for the reference's
 * erexit (neither cut requested). */
int orc_f_bandpass(float flo, float fhi, int nplo, int nphi, int phase, float dt, int nt, int ntr, float *tr)
{
    const float pi = 3.1415926535f;
    const float flo_t = flo * dt, fhi_t = fhi * dt;
    const int npad = 10, n1 = nt + 2 * npad;
    if (flo_t < 0.0001f && fhi_t > 0.4999f) return -1;
    float *data = (float *)malloc(sizeof(float) * n1), *nw = (float *)malloc(sizeof(float) * n1),
          *tmp = (float *)malloc(sizeof(float) * n1);
    for (int k = 0; k < ntr; k++) {
        float *in = tr + (size_t)k * nt;
        for (int i = 0; i < n1; i++) data[i] = nw[i] = tmp[i] = 0.f;
        for (int i = 0; i < nt; i++) data[npad + i] = in[i];
        for (int j = 1; j <= npad; j++) {
            float s1 = sinf((float)(j - 1) * pi / (float)(2 * (npad - 1)));
            float s2 = sinf((float)(npad - j) * pi / (float)(2 * (npad - 1)));
            data[j - 1] = in[0] * (s1 * s1);
            data[n1 - npad + j - 1] = in[nt - 1] * (s2 * s2);
        }
        if (flo_t > 0.000001f) orc_lowcut(flo_t, nplo, phase, n1, data, nw, tmp);
        if (fhi_t < 4.999999f) orc_highcut(fhi_t, nphi, phase, n1, data, nw, tmp);
        for (int i = 0; i < nt; i++) in[i] = data[npad + i];
    }
    free(data);
    free(nw);
    free(tmp);
    return 0;
}
