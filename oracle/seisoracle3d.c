/*
 * seisoracle3d.c — CPU oracle for 3D forward modeling (FP32).
 *
 * TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker, never by the
 * product.  It restates the reference's 3D constant-density forward modeling
 * (Acoustic_modeling_sep.f90:112-121 with genpar%twoD false, nbound 4):
 *   window and padding  ExtractPadModel.f90:19-85 (patch with apertures in x
 *                       and y), :183-186 (nxw, nyw), :509-530 mod_window_pad,
 *                       ReadSouVelRho.f90:545-596 model_pad
 *   Higdon tables       Boundary.f90:17-63 (gz per (x,y), gx per (z,y), gy per (z,x))
 *   stencil             FDderivatives.f90:97-131 FD_2nd_3D_derivatives_scalar_forward_grid
 *   injection           Injection.f90:151-194 Injection_sinc -> :274-346
 *                       Injection_source_sinc_xyz (y weights from the x sinc,
 *                       sinc(:,2), as the reference does)
 *   extraction          Extraction.f90:109-134 (sinc, gain v, not v^2, in 3D),
 *                       :71-106 (nearest), :217-246 (LSRTM, 1/(dz dx dy)),
 *                       :137-162 (sinc, gain 1) -> :276-337 Extraction_1trace_sinc_xyz
 *                       (y weights from the x sinc), summed k, j, i in order
 *   time update         FDderivatives.f90:960-982
 *   Higdon bands        Boundary.f90:89-184 Boundary_3d (surf_type 0)
 * Pinned bitwise against the compiled reference kernels composed in
 * Propagator.f90 order (tests/test_oracle_3d.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "seisfd.h"

void orc_f_mksinc(float *sinc, int lsinc, float shift);
void orc_f_higdon_coeff(float nu, float beta1, float beta2, float weight, float *c);

typedef struct {
    int nmin1, nmax1, nmin2, nmax2, nmin3, nmax3;
    int NZ, NX, NY;
} B3;

#define IX3(b, i, j, k)                                                                                       \
    ((size_t)((i) - ((b)->nmin1 - 4)) + (size_t)((j) - ((b)->nmin2 - 4)) * (size_t)(b)->NZ +                 \
     (size_t)((k) - ((b)->nmin3 - 4)) * (size_t)(b)->NZ * (size_t)(b)->NX)

static int nint3(float x) { return (int)roundf(x); }

static void find_i_d3(int *ic, float *dc, float c, float o, float d)
{
    *ic = nint3((c - o) / d) + 1;
    *dc = c - (o + (float)(*ic - 1) * d);
}

/* extract_coord_source_receiver_patch (x and y) and nxw / nyw */
void orc_window3d(const sfd_grid3d *g, const sfd_shot3d *s, float *oxw, int *nxw, float *oyw, int *nyw)
{
    float sx = s->sx, sy = s->sy, gx1 = s->rx[0], gxn = s->rx[0], gy1 = s->ry[0], gyn = s->ry[0];
    for (int i = 1; i < s->ntr; i++) {
        gx1 = gx1 < s->rx[i] ? gx1 : s->rx[i];
        gxn = gxn > s->rx[i] ? gxn : s->rx[i];
        gy1 = gy1 < s->ry[i] ? gy1 : s->ry[i];
        gyn = gyn > s->ry[i] ? gyn : s->ry[i];
    }
    float minx = sx < gx1 ? sx : gx1, miny = sy < gy1 ? sy : gy1;
    float maxx = sx > gxn ? sx : gxn, maxy = sy > gyn ? sy : gyn;
    float mx1 = (gx1 + sx) / 2, mxn = (gxn + sx) / 2, my1 = (gy1 + sy) / 2, myn = (gyn + sy) / 2;
    float cx = mx1 + fabsf(mxn - mx1) / 2, cy = my1 + fabsf(myn - my1) / 2;
    *oxw = minx < cx - g->aperture_x / 2 ? minx : cx - g->aperture_x / 2;
    float endx = maxx > cx + g->aperture_x / 2 ? maxx : cx + g->aperture_x / 2;
    *oyw = miny < cy - g->aperture_y / 2 ? miny : cy - g->aperture_y / 2;
    float endy = maxy > cy + g->aperture_y / 2 ? maxy : cy + g->aperture_y / 2;
    *nxw = nint3((endx - *oxw) / g->dx) + 1;
    *nyw = nint3((endy - *oyw) / g->dy) + 1;
}

void orc_bounds3d(B3 *b, int nz, int nxw, int nyw, int ntaper)
{
    b->nmin1 = 1 - ntaper; b->nmax1 = nz + ntaper;
    b->nmin2 = 1 - ntaper; b->nmax2 = nxw + ntaper;
    b->nmin3 = 1 - ntaper; b->nmax3 = nyw + ntaper;
    b->NZ = b->nmax1 - b->nmin1 + 9;
    b->NX = b->nmax2 - b->nmin2 + 9;
    b->NY = b->nmax3 - b->nmin3 + 9;
}

/* mod_window_pad (nearest model column and line, clamped) then model_pad
 * (edge replication): padded(i,j,k) = the window's value at the clamped
 * (i,j,k) */
void orc_window_pad3d(const sfd_grid3d *g, const float *big, float oxw, int nxw, float oyw, int nyw, const B3 *b,
                      float *out)
{
    for (int k = b->nmin3 - 4; k <= b->nmax3 + 4; k++) {
        int kw = k < 1 ? 1 : (k > nyw ? nyw : k);
        float cyy = (float)(kw - 1) * g->dy + oyw;
        int iy = nint3((cyy - g->oy) / g->dy) + 1;
        iy = iy < 1 ? 1 : (iy > g->ny ? g->ny : iy);
        for (int j = b->nmin2 - 4; j <= b->nmax2 + 4; j++) {
            int jw = j < 1 ? 1 : (j > nxw ? nxw : j);
            float cxx = (float)(jw - 1) * g->dx + oxw;
            int ix = nint3((cxx - g->ox) / g->dx) + 1;
            ix = ix < 1 ? 1 : (ix > g->nx ? g->nx : ix);
            for (int i = b->nmin1 - 4; i <= b->nmax1 + 4; i++) {
                int iw = i < 1 ? 1 : (i > g->nz ? g->nz : i);
                out[IX3(b, i, j, k)] =
                    (float)0 + big[(size_t)(iw - 1) + (size_t)(ix - 1) * g->nz + (size_t)(iy - 1) * g->nz * g->nx];
            }
        }
    }
}

typedef struct {
    B3 b;
    float dz, dx, dy, dt;
    float *vel;
    float *gz, *gx, *gy;   /* Higdon tables: 16 per (j,k), (i,k), (i,j) of the FD region */
    float c[15];           /* c0z..c4z, c0x..c4x, c0y..c4y */
    float *rho, *delp;     /* variable density: padded rho and 1/rho2 (NULL: constant density) */
    float *sz, *sx, *sy;   /* staggered gradients s3dzz, s3dxx, s3dyy (cells never written stay 0) */
} P3;

#define A3(P, u, i, j, k) (u)[IX3(&(P)->b, i, j, k)]

static void higdon3(P3 *P)
{
    const B3 *b = &P->b;
    const int n1 = b->nmax1 - b->nmin1 + 1, n2 = b->nmax2 - b->nmin2 + 1, n3 = b->nmax3 - b->nmin3 + 1;
    const float dxi = 1.f / P->dx, dyi = 1.f / P->dy, dzi = 1.f / P->dz;
    P->gz = (float *)calloc((size_t)16 * n2 * n3, sizeof(float));
    P->gx = (float *)calloc((size_t)16 * n1 * n3, sizeof(float));
    P->gy = (float *)calloc((size_t)16 * n1 * n2, sizeof(float));
    for (int k = b->nmin3; k <= b->nmax3; k++)
        for (int j = b->nmin2; j <= b->nmax2; j++) {
            float *g = P->gz + 16 * ((size_t)(j - b->nmin2) + (size_t)(k - b->nmin3) * n2);
            orc_f_higdon_coeff(A3(P, P->vel, b->nmin1, j, k) * P->dt * dzi, 1.0f, 0.7f, 0.5f, g);
            orc_f_higdon_coeff(A3(P, P->vel, b->nmax1, j, k) * P->dt * dzi, 1.0f, 0.7f, 0.5f, g + 8);
        }
    for (int k = b->nmin3; k <= b->nmax3; k++)
        for (int i = b->nmin1; i <= b->nmax1; i++) {
            float *g = P->gx + 16 * ((size_t)(i - b->nmin1) + (size_t)(k - b->nmin3) * n1);
            orc_f_higdon_coeff(A3(P, P->vel, i, b->nmin2, k) * P->dt * dxi, 1.0f, 0.7f, 0.5f, g);
            orc_f_higdon_coeff(A3(P, P->vel, i, b->nmax2, k) * P->dt * dxi, 1.0f, 0.7f, 0.5f, g + 8);
        }
    for (int j = b->nmin2; j <= b->nmax2; j++)
        for (int i = b->nmin1; i <= b->nmax1; i++) {
            float *g = P->gy + 16 * ((size_t)(i - b->nmin1) + (size_t)(j - b->nmin2) * n1);
            orc_f_higdon_coeff(A3(P, P->vel, i, j, b->nmin3) * P->dt * dyi, 1.0f, 0.7f, 0.5f, g);
            orc_f_higdon_coeff(A3(P, P->vel, i, j, b->nmax3) * P->dt * dyi, 1.0f, 0.7f, 0.5f, g + 8);
        }
}

static void stencil3(P3 *P, const float *u2, float *u3)
{
    const B3 *b = &P->b;
    const float *c = P->c;
#pragma omp parallel for schedule(static)
    for (int k = b->nmin3; k <= b->nmax3; k++)
        for (int j = b->nmin2; j <= b->nmax2; j++)
            for (int i = b->nmin1; i <= b->nmax1; i++) {
#define U(ii, jj, kk) A3(P, u2, ii, jj, kk)
                float tz = c[0] * (U(i, j, k)) + c[1] * (U(i + 1, j, k) + U(i - 1, j, k)) +
                           c[2] * (U(i + 2, j, k) + U(i - 2, j, k)) + c[3] * (U(i + 3, j, k) + U(i - 3, j, k)) +
                           c[4] * (U(i + 4, j, k) + U(i - 4, j, k));
                float tx = c[5] * (U(i, j, k)) + c[6] * (U(i, j + 1, k) + U(i, j - 1, k)) +
                           c[7] * (U(i, j + 2, k) + U(i, j - 2, k)) + c[8] * (U(i, j + 3, k) + U(i, j - 3, k)) +
                           c[9] * (U(i, j + 4, k) + U(i, j - 4, k));
                float ty = c[10] * (U(i, j, k)) + c[11] * (U(i, j, k + 1) + U(i, j, k - 1)) +
                           c[12] * (U(i, j, k + 2) + U(i, j, k - 2)) + c[13] * (U(i, j, k + 3) + U(i, j, k - 3)) +
                           c[14] * (U(i, j, k + 4) + U(i, j, k - 4));
#undef U
                float v = A3(P, P->vel, i, j, k);
                A3(P, u3, i, j, k) = v * v * (tx + tz + ty);
            }
}

/* FDderivatives.f90:263-382 FD_3D_derivatives_acoustic_forward_grid:
 * staggered gradients times 1/rho2 on the FD region; first-order gradients
 * times the boundary's 1/rho2 in the 3 operator-padding layers of each face
 * (top/bottom for (x, y) in the FD region, left/right for (z, y), front/back
 * for (z, x)); the outermost layer and the edges and corners are never
 * written and keep their initial zero; then the staggered divergence times
 * v^2 rho, summed x, y, z. */
static void stencil3_rho(P3 *P, const float *u2, float *u3)
{
    const B3 *b = &P->b;
    const float *c = P->c;
    const float dzi = 1.f / P->dz, dxi = 1.f / P->dx, dyi = 1.f / P->dy;
    float *sz = P->sz, *sx = P->sx, *sy = P->sy;
#define U(ii, jj, kk) A3(P, u2, ii, jj, kk)
#define D(ii, jj, kk) A3(P, P->delp, ii, jj, kk)
#define FO(i, j, k, dd)                                                                             \
    do {                                                                                            \
        A3(P, sz, i, j, k) = dzi * (U(i + 1, j, k) - U(i, j, k)) * (dd);                            \
        A3(P, sx, i, j, k) = dxi * (U(i, j + 1, k) - U(i, j, k)) * (dd);                            \
        A3(P, sy, i, j, k) = dyi * (U(i, j, k + 1) - U(i, j, k)) * (dd);                            \
    } while (0)
#pragma omp parallel for schedule(static)
    for (int k = b->nmin3; k <= b->nmax3; k++)
        for (int j = b->nmin2; j <= b->nmax2; j++)
            for (int i = b->nmin1; i <= b->nmax1; i++) {
                A3(P, sz, i, j, k) = (c[1] * (U(i + 1, j, k) - U(i, j, k)) + c[2] * (U(i + 2, j, k) - U(i - 1, j, k)) +
                                      c[3] * (U(i + 3, j, k) - U(i - 2, j, k)) + c[4] * (U(i + 4, j, k) - U(i - 3, j, k))) *
                                     D(i, j, k);
                A3(P, sx, i, j, k) = (c[6] * (U(i, j + 1, k) - U(i, j, k)) + c[7] * (U(i, j + 2, k) - U(i, j - 1, k)) +
                                      c[8] * (U(i, j + 3, k) - U(i, j - 2, k)) + c[9] * (U(i, j + 4, k) - U(i, j - 3, k))) *
                                     D(i, j, k);
                A3(P, sy, i, j, k) = (c[11] * (U(i, j, k + 1) - U(i, j, k)) + c[12] * (U(i, j, k + 2) - U(i, j, k - 1)) +
                                      c[13] * (U(i, j, k + 3) - U(i, j, k - 2)) + c[14] * (U(i, j, k + 4) - U(i, j, k - 3))) *
                                     D(i, j, k);
            }
    for (int k = b->nmin3; k <= b->nmax3; k++) {
        for (int j = b->nmin2; j <= b->nmax2; j++) {
            for (int i = b->nmin1 - 1; i >= b->nmin1 - 3; i--) FO(i, j, k, D(b->nmin1, j, k));
            for (int i = b->nmax1 + 1; i <= b->nmax1 + 3; i++) FO(i, j, k, D(b->nmax1, j, k));
        }
        for (int j = b->nmin2 - 1; j >= b->nmin2 - 3; j--)
            for (int i = b->nmin1; i <= b->nmax1; i++) FO(i, j, k, D(i, b->nmin2, k));
        for (int j = b->nmax2 + 1; j <= b->nmax2 + 3; j++)
            for (int i = b->nmin1; i <= b->nmax1; i++) FO(i, j, k, D(i, b->nmax2, k));
    }
    for (int k = b->nmin3 - 1; k >= b->nmin3 - 3; k--)
        for (int j = b->nmin2; j <= b->nmax2; j++)
            for (int i = b->nmin1; i <= b->nmax1; i++) FO(i, j, k, D(i, j, b->nmin3));
    for (int k = b->nmax3 + 1; k <= b->nmax3 + 3; k++)
        for (int j = b->nmin2; j <= b->nmax2; j++)
            for (int i = b->nmin1; i <= b->nmax1; i++) FO(i, j, k, D(i, j, b->nmax3));
#undef FO
#undef D
#undef U
#define S(a, ii, jj, kk) A3(P, a, ii, jj, kk)
#pragma omp parallel for schedule(static)
    for (int k = b->nmin3; k <= b->nmax3; k++)
        for (int j = b->nmin2; j <= b->nmax2; j++)
            for (int i = b->nmin1; i <= b->nmax1; i++) {
                float tz = c[1] * (S(sz, i, j, k) - S(sz, i - 1, j, k)) + c[2] * (S(sz, i + 1, j, k) - S(sz, i - 2, j, k)) +
                           c[3] * (S(sz, i + 2, j, k) - S(sz, i - 3, j, k)) + c[4] * (S(sz, i + 3, j, k) - S(sz, i - 4, j, k));
                float tx = c[6] * (S(sx, i, j, k) - S(sx, i, j - 1, k)) + c[7] * (S(sx, i, j + 1, k) - S(sx, i, j - 2, k)) +
                           c[8] * (S(sx, i, j + 2, k) - S(sx, i, j - 3, k)) + c[9] * (S(sx, i, j + 3, k) - S(sx, i, j - 4, k));
                float ty = c[11] * (S(sy, i, j, k) - S(sy, i, j, k - 1)) + c[12] * (S(sy, i, j, k + 1) - S(sy, i, j, k - 2)) +
                           c[13] * (S(sy, i, j, k + 2) - S(sy, i, j, k - 3)) + c[14] * (S(sy, i, j, k + 3) - S(sy, i, j, k - 4));
                float v = A3(P, P->vel, i, j, k);
                A3(P, u3, i, j, k) = v * v * A3(P, P->rho, i, j, k) * (tx + ty + tz);
            }
#undef S
}

/* Interpolate.f90:44-130 in 3D (nbound 4): rho2 = rho sinc-interpolated half
 * a cell along z, then x, then y (9-point sinc at shift 0.5, each line
 * extended by its end values); delp = 1/rho2 (allocate_sxx3d_syy3d_szz3d_delp3d,
 * FDderivatives.f90:385-395). */
void orc_rho_half3d(const B3 *b, const float *rho, float *delp)
{
    const int L = 9, h = L / 2;
    float sinc[9];
    orc_f_mksinc(sinc, L, 0.5f);
    const int n[3] = {b->NZ, b->NX, b->NY};
    const size_t st[3] = {1, (size_t)b->NZ, (size_t)b->NZ * b->NX};
    const size_t ntot = (size_t)b->NZ * b->NX * b->NY;
    float *cur = (float *)malloc(sizeof(float) * ntot), *t = (float *)malloc(sizeof(float) * (size_t)(b->NZ + b->NX + b->NY + 2 * L));
    memcpy(cur, rho, sizeof(float) * ntot);
    for (int ax = 0; ax < 3; ax++) {
        const int a1 = (ax + 1) % 3, a2 = (ax + 2) % 3, m = n[ax];
        for (int p = 0; p < n[a2]; p++)
            for (int q = 0; q < n[a1]; q++) {
                const size_t base = (size_t)p * st[a2] + (size_t)q * st[a1];
                float *tt = t + L;   /* tt[i] for i in -L .. m-1+L */
                for (int i = 0; i < m; i++) tt[i] = cur[base + (size_t)i * st[ax]];
                for (int i = -L; i < 0; i++) tt[i] = tt[0];
                for (int i = m; i < m + L; i++) tt[i] = tt[m - 1];
                for (int i = 0; i < m; i++) {
                    float f = 0;
                    for (int l = -h; l <= h; l++) f = f + sinc[h + l] * tt[i + l];
                    cur[base + (size_t)i * st[ax]] = f;
                }
            }
    }
    for (size_t q = 0; q < ntot; q++) delp[q] = 1.f / cur[q];
    free(cur);
    free(t);
}

static void update3(P3 *P, const float *um, const float *u0, const float *u1, const float *u2, float *u3)
{
    const B3 *b = &P->b;
    const float div11 = 1.f / 11.f, dt2 = P->dt * P->dt;
#pragma omp parallel for schedule(static)
    for (int k = b->nmin3; k <= b->nmax3; k++)
        for (int j = b->nmin2; j <= b->nmax2; j++)
            for (int i = b->nmin1; i <= b->nmax1; i++) {
                size_t q = IX3(b, i, j, k);
                u3[q] = (20.f * u2[q] - 6.f * u1[q] - 4.f * u0[q] + um[q] + 12.f * dt2 * u3[q]) * div11;
            }
}

/* Higdon_factor with (u0, u1, u2) = (grid%u1, grid%u2, grid%u3) */
static void hfac3(P3 *P, const float *U1, const float *U2, float *U3, int i, int j, int k, int axis)
{
    const B3 *b = &P->b;
    const int n1 = b->nmax1 - b->nmin1 + 1, n2 = b->nmax2 - b->nmin2 + 1;
    const float *g;
    int di = 0, dj = 0, dk = 0;
    switch (axis) {
    case -1: g = P->gz + 16 * ((size_t)(j - b->nmin2) + (size_t)(k - b->nmin3) * n2); di = 1; break;
    case 1: g = P->gz + 16 * ((size_t)(j - b->nmin2) + (size_t)(k - b->nmin3) * n2) + 8; di = -1; break;
    case -2: g = P->gx + 16 * ((size_t)(i - b->nmin1) + (size_t)(k - b->nmin3) * n1); dj = 1; break;
    case 2: g = P->gx + 16 * ((size_t)(i - b->nmin1) + (size_t)(k - b->nmin3) * n1) + 8; dj = -1; break;
    case -3: g = P->gy + 16 * ((size_t)(i - b->nmin1) + (size_t)(j - b->nmin2) * n1); dk = 1; break;
    default: g = P->gy + 16 * ((size_t)(i - b->nmin1) + (size_t)(j - b->nmin2) * n1) + 8; dk = -1; break;
    }
    float f = g[0] * A3(P, U3, i + di, j + dj, k + dk) + g[1] * A3(P, U3, i + 2 * di, j + 2 * dj, k + 2 * dk) +
              g[2] * A3(P, U2, i, j, k) + g[3] * A3(P, U2, i + di, j + dj, k + dk) +
              g[4] * A3(P, U2, i + 2 * di, j + 2 * dj, k + 2 * dk) + g[5] * A3(P, U1, i, j, k) +
              g[6] * A3(P, U1, i + di, j + dj, k + dk) + g[7] * A3(P, U1, i + 2 * di, j + 2 * dj, k + 2 * dk);
    A3(P, U3, i, j, k) = -f;
}

static void boundary3(P3 *P, const float *u1, const float *u2, float *u3)
{
    const B3 *b = &P->b;
    const int nz0 = b->nmin1;
    for (int k = b->nmin3; k <= b->nmax3; k++) {
        for (int j = b->nmin2; j <= b->nmax2; j++) {
            for (int i = nz0 + 3; i >= b->nmin1 - 4; i--) hfac3(P, u1, u2, u3, i, j, k, -1);
            for (int i = b->nmax1 - 3; i <= b->nmax1 + 4; i++) hfac3(P, u1, u2, u3, i, j, k, 1);
        }
        for (int j = b->nmin2 + 3; j >= b->nmin2 - 4; j--)
            for (int i = nz0; i <= b->nmax1; i++) hfac3(P, u1, u2, u3, i, j, k, -2);
        for (int j = b->nmax2 - 3; j <= b->nmax2 + 4; j++)
            for (int i = nz0; i <= b->nmax1; i++) hfac3(P, u1, u2, u3, i, j, k, 2);
    }
    for (int j = b->nmin2; j <= b->nmax2; j++) {
        for (int k = b->nmin3 + 3; k >= b->nmin3 - 4; k--)
            for (int i = nz0; i <= b->nmax1; i++) hfac3(P, u1, u2, u3, i, j, k, -3);
        for (int k = b->nmax3 - 3; k <= b->nmax3 + 4; k++)
            for (int i = nz0; i <= b->nmax1; i++) hfac3(P, u1, u2, u3, i, j, k, 3);
    }
}

void orc_f_half_step(float *tr, int nt, int lsinc);

/* 3D forward modeling of one shot: traces (nt, ntr) time fastest.  rho_full
 * (nz, nx, ny) or NULL: with a density, read_window_vel's rho (windowed and
 * padded like vel, then Interpolate), coefficients over delta
 * (FD_acoustic_rho_init_coefs, coefpower 1), the density stencil, the v^2 rho
 * injection gain (Injection.f90:181-187) and readsou's half-step source shift
 * (ReadSouVelRho.f90:328-347). */
/* The shot's padded model (window, vel, with a density rho and 1/rho2),
 * coefficients and Higdon tables. */
static int setup3(P3 *P, const sfd_grid3d *g, const float *vel_full, const float *rho_full, const sfd_shot3d *s,
                  float *oxw, int *nxw, float *oyw, int *nyw)
{
    memset(P, 0, sizeof *P);
    orc_window3d(g, s, oxw, nxw, oyw, nyw);
    orc_bounds3d(&P->b, g->nz, *nxw, *nyw, g->ntaper);
    const B3 *b = &P->b;
    const size_t n = (size_t)b->NZ * b->NX * b->NY;
    P->dz = g->dz; P->dx = g->dx; P->dy = g->dy; P->dt = g->dt;
    P->vel = (float *)calloc(n, sizeof(float));
    if (!P->vel) return -2;
    orc_window_pad3d(g, vel_full, *oxw, *nxw, *oyw, *nyw, b, P->vel);
    const float dzi = 1.f / g->dz, dxi = 1.f / g->dx, dyi = 1.f / g->dy;
    if (rho_full) {
        P->rho = (float *)calloc(n, sizeof(float));
        P->delp = (float *)calloc(n, sizeof(float));
        P->sz = (float *)calloc(n, sizeof(float));
        P->sx = (float *)calloc(n, sizeof(float));
        P->sy = (float *)calloc(n, sizeof(float));
        if (!P->sy) return -2;
        orc_window_pad3d(g, rho_full, *oxw, *nxw, *oyw, *nyw, b, P->rho);
        orc_rho_half3d(b, P->rho, P->delp);
        /* FDcoefs.f90:12-23 over delta per axis (coefpower 1) */
        const float r[5] = {0.f, 1225.f / 1024.f, -(1225.f / (1024.f * 15.f)), 1225.f / (1024.f * 125.f),
                            -(1225.f / (1024.f * 1715.f))};
        for (int k = 0; k < 5; k++) {
            P->c[k] = r[k] * dzi;
            P->c[5 + k] = r[k] * dxi;
            P->c[10 + k] = r[k] * dyi;
        }
    } else {
        /* FDcoefs.f90:25-36 over delta^2 per axis */
        const float cc[5] = {-205.f / 72.f, 8.f / 5, -1.f / 5.f, 8.f / 315.f, -1.f / 560.f};
        for (int k = 0; k < 5; k++) {
            P->c[k] = cc[k] * (dzi * dzi);
            P->c[5 + k] = cc[k] * (dxi * dxi);
            P->c[10 + k] = cc[k] * (dyi * dyi);
        }
    }
    higdon3(P);
    return 0;
}

static void free3(P3 *P)
{
    free(P->vel); free(P->gz); free(P->gx); free(P->gy);
    free(P->rho); free(P->delp); free(P->sz); free(P->sx); free(P->sy);
}

int orc_f_modeling3d_rho(const sfd_grid3d *g, const float *vel_full, const float *rho_full, const sfd_shot3d *s,
                         float *traces, int extract_kind)
{
    P3 P;
    float oxw, oyw;
    int nxw, nyw;
    if (setup3(&P, g, vel_full, rho_full, s, &oxw, &nxw, &oyw, &nyw)) return -2;
    const B3 *b = &P.b;
    const size_t n = (size_t)b->NZ * b->NX * b->NY;
    float *buf[5];
    for (int q = 0; q < 5; q++) buf[q] = (float *)calloc(n, sizeof(float));
    if (!buf[4]) return -2;
    float *wav = (float *)malloc(sizeof(float) * (size_t)s->nt);
    memcpy(wav, s->wavelet, sizeof(float) * (size_t)s->nt);
    if (rho_full) orc_f_half_step(wav, s->nt, g->lsinc);
    const int l = g->lsinc, h = (int)(l * 0.5);
    /* source */
    int sic[3];
    float sdc[3], sw[3][64];
    find_i_d3(&sic[0], &sdc[0], s->sz, g->oz, g->dz);
    find_i_d3(&sic[1], &sdc[1], s->sx, oxw, g->dx);
    find_i_d3(&sic[2], &sdc[2], s->sy, oyw, g->dy);
    orc_f_mksinc(sw[0], l, sdc[0] / g->dz);
    orc_f_mksinc(sw[1], l, sdc[1] / g->dx);
    orc_f_mksinc(sw[2], l, sdc[2] / g->dy);
    /* receivers */
    int *ric = (int *)malloc(sizeof(int) * 3 * (size_t)s->ntr);
    float *rw = (float *)malloc(sizeof(float) * 2 * 64 * (size_t)s->ntr);
    for (int r = 0; r < s->ntr; r++) {
        float dcz, dcx, dcy;
        find_i_d3(&ric[3 * r], &dcz, s->rz[r], g->oz, g->dz);
        find_i_d3(&ric[3 * r + 1], &dcx, s->rx[r], oxw, g->dx);
        find_i_d3(&ric[3 * r + 2], &dcy, s->ry[r], oyw, g->dy);
        orc_f_mksinc(rw + 128 * r, l, dcz / g->dz);
        orc_f_mksinc(rw + 128 * r + 64, l, dcx / g->dx);
        (void)dcy;   /* the y weights are the x sinc (Extraction.f90:315) */
    }
    memset(traces, 0, sizeof(float) * (size_t)s->nt * s->ntr);
    for (int it = 1; it <= s->nt; it++) {
        float *u2 = buf[3], *u3 = buf[4];
        /* ExtractData */
        for (int r = 0; r < s->ntr; r++) {
            const int *ic = ric + 3 * r;
            float *tr = traces + (size_t)r * s->nt + (it - 1);
            if (extract_kind == SFD_EXTRACT_NEAREST) {
                *tr = *tr + A3(&P, u2, ic[0], ic[1], ic[2]);
                continue;
            }
            float gain = extract_kind == SFD_EXTRACT_SINC_V2 ? A3(&P, P.vel, ic[0], ic[1], ic[2])
                         : extract_kind == SFD_EXTRACT_SINC_LSRTM ? 1.f / (g->dz * g->dx * g->dy)
                                                                  : 1.f;
            const float *wz = rw + 128 * r, *wx = wz + 64;
            float acc = *tr;
            for (int k = -h; k <= h; k++)
                for (int j = -h; j <= h; j++)
                    for (int i = -h; i <= h; i++)
                        acc = acc + gain * A3(&P, u2, i + ic[0], j + ic[1], k + ic[2]) * wz[h + i] * wx[h + j] * wx[h + k];
            *tr = acc;
        }
        if (P.rho) stencil3_rho(&P, u2, u3);
        else stencil3(&P, u2, u3);
        /* Injection_sinc, v2r = vel^2 (times rho with a density) */
        {
            float v = A3(&P, P.vel, sic[0], sic[1], sic[2]), v2r = v * v, tr = wav[it - 1];
            if (P.rho) v2r = v2r * A3(&P, P.rho, sic[0], sic[1], sic[2]);
            for (int k = -h; k <= h; k++)
                for (int j = -h; j <= h; j++)
                    for (int i = -h; i <= h; i++)
                        A3(&P, u3, i + sic[0], j + sic[1], k + sic[2]) =
                            A3(&P, u3, i + sic[0], j + sic[1], k + sic[2]) + v2r * tr * sw[0][h + i] * sw[1][h + j] * sw[1][h + k];
        }
        update3(&P, buf[0], buf[1], buf[2], buf[3], buf[4]);
        boundary3(&P, buf[2], buf[3], buf[4]);
        float *t = buf[0];
        buf[0] = buf[1]; buf[1] = buf[2]; buf[2] = buf[3]; buf[3] = buf[4]; buf[4] = t;
    }
    for (int q = 0; q < 5; q++) free(buf[q]);
    free(ric); free(rw); free(wav);
    free3(&P);
    return 0;
}

int orc_f_modeling3d(const sfd_grid3d *g, const float *vel_full, const sfd_shot3d *s, float *traces, int extract_kind)
{
    return orc_f_modeling3d_rho(g, vel_full, NULL, s, traces, extract_kind);
}

void orc_f_compute_taper(float *t, int n);

/* Injection_sinc (Injection.f90:151-194 -> :274-346) of ntr traces at their
 * positions, in trace order (the harness runs the reference's OpenMP loop
 * over traces on one thread): u += (((v^2 [rho]) s wz) wx) wy, y weights from
 * the x sinc. */
typedef struct { int ic[3]; float wz[64], wx[64]; } Pt3;

static void inject3(P3 *P, float *u, const Pt3 *pt, int ntr, const float *tr, int nt, int it, int h)
{
    for (int r = 0; r < ntr; r++) {
        const int *ic = pt[r].ic;
        float v = A3(P, P->vel, ic[0], ic[1], ic[2]), v2r = v * v;
        if (P->rho) v2r = v2r * A3(P, P->rho, ic[0], ic[1], ic[2]);
        const float s = tr[(size_t)r * nt + (it - 1)];
        for (int k = -h; k <= h; k++)
            for (int j = -h; j <= h; j++)
                for (int i = -h; i <= h; i++)
                    A3(P, u, i + ic[0], j + ic[1], k + ic[2]) =
                        A3(P, u, i + ic[0], j + ic[1], k + ic[2]) + v2r * s * pt[r].wz[h + i] * pt[r].wx[h + j] *
                                                                        pt[r].wx[h + k];
    }
}

static void points3(const sfd_grid3d *g, float oxw, float oyw, int n, const float *z, const float *x, const float *y,
                    Pt3 *pt)
{
    for (int r = 0; r < n; r++) {
        float dc[3];
        find_i_d3(&pt[r].ic[0], &dc[0], z[r], g->oz, g->dz);
        find_i_d3(&pt[r].ic[1], &dc[1], x[r], oxw, g->dx);
        find_i_d3(&pt[r].ic[2], &dc[2], y[r], oyw, g->dy);
        orc_f_mksinc(pt[r].wz, g->lsinc, dc[0] / g->dz);
        orc_f_mksinc(pt[r].wx, g->lsinc, dc[1] / g->dx);
    }
}

/* One time step after extraction / imaging: stencil, injection, update,
 * Higdon, rotation. */
static void step3(P3 *P, float **buf, const Pt3 *pt, int npt, const float *tr, int nt, int it, int h)
{
    if (P->rho) stencil3_rho(P, buf[3], buf[4]);
    else stencil3(P, buf[3], buf[4]);
    inject3(P, buf[4], pt, npt, tr, nt, it, h);
    update3(P, buf[0], buf[1], buf[2], buf[3], buf[4]);
    boundary3(P, buf[2], buf[3], buf[4]);
    float *t = buf[0];
    buf[0] = buf[1]; buf[1] = buf[2]; buf[2] = buf[3]; buf[3] = buf[4]; buf[4] = t;
}

/* progs/rtm/3DRTM.f90 -> RTM_to_memory (to_disk_to_memory_RTM.f90:358-492)
 * with genpar%twoD false: the forward pass stores u2 on the window every
 * snapi steps (Extraction_wavefield, Extraction.f90:562-601), the backward
 * pass runs the same forward stencil (the 3D branch's
 * FD_2nd_3D_derivatives_scalar_forward_grid, or the density operator) on
 * vel_bwd when given (mod%vel2, :425) with its own Higdon tables, injects the
 * receivers' traces by Injection_sinc and images with
 * Imaging_condition_sourceonly_from_memory (Imaging.f90:228-275) before the
 * stencil of every snapi-th step; compute_taper (Taper.f90:14-28) in x and
 * y; mod_copy_image (ExtractPadModel.f90:662-695) adds the window into
 * image / illum (nz, nx, ny).  Optional outputs: the window image and
 * illumination, the snapshots and the backward u2 at each imaging step
 * (nz*nxw*nyw each, in the order they are taken). */
int orc_f_rtm3d(const sfd_grid3d *g, const float *vel_full, const float *vel_bwd, const float *rho_full,
                const sfd_shot3d *s, const float *traces_in, float *image, float *illum, float *win_image,
                float *win_illum, float *snaps_out, float *capture)
{
    P3 P;
    float oxw, oyw;
    int nxw, nyw;
    if (g->snapi < 1) return -1;
    if (setup3(&P, g, vel_full, rho_full, s, &oxw, &nxw, &oyw, &nyw)) return -2;
    const B3 *b = &P.b;
    const size_t n = (size_t)b->NZ * b->NX * b->NY, nwin = (size_t)g->nz * nxw * nyw;
    const int ntsnap = s->nt / g->snapi, h = (int)(g->lsinc * 0.5);
    float *buf[5], *snaps = (float *)calloc(nwin * (ntsnap > 0 ? ntsnap : 1), sizeof(float));
    float *img = (float *)calloc(nwin, sizeof(float)), *ill = (float *)calloc(nwin, sizeof(float));
    float *tx = (float *)malloc(sizeof(float) * nxw), *ty = (float *)malloc(sizeof(float) * nyw);
    for (int q = 0; q < 5; q++) buf[q] = (float *)calloc(n, sizeof(float));
    if (!buf[4] || !snaps || !img || !ill || !tx || !ty) return -2;
    float *wav = (float *)malloc(sizeof(float) * (size_t)s->nt);
    memcpy(wav, s->wavelet, sizeof(float) * (size_t)s->nt);
    if (rho_full) orc_f_half_step(wav, s->nt, g->lsinc);
    Pt3 src, *rec = (Pt3 *)malloc(sizeof(Pt3) * (size_t)(s->ntr > 0 ? s->ntr : 1));
    points3(g, oxw, oyw, 1, &s->sz, &s->sx, &s->sy, &src);
    points3(g, oxw, oyw, s->ntr, s->rz, s->rx, s->ry, rec);
#define WIN(u, i, j, k) A3(&P, u, i, j, k)
    int counter = 0;
    for (int it = 1; it <= s->nt; it++) {
        if (it % g->snapi == 0) {   /* Extraction_wavefield */
            float *S = snaps + (size_t)counter * nwin;
            counter++;
            for (int k = 1; k <= nyw; k++)
                for (int j = 1; j <= nxw; j++)
                    for (int i = 1; i <= g->nz; i++)
                        S[(i - 1) + (size_t)(j - 1) * g->nz + (size_t)(k - 1) * g->nz * nxw] = WIN(buf[3], i, j, k);
        }
        step3(&P, buf, &src, 1, wav, s->nt, it, h);
    }
    /* backward: a new propagator call (zeroed levels and gradient arrays) */
    for (int q = 0; q < 5; q++) memset(buf[q], 0, n * sizeof(float));
    if (P.sz) {
        memset(P.sz, 0, n * sizeof(float));
        memset(P.sx, 0, n * sizeof(float));
        memset(P.sy, 0, n * sizeof(float));
    }
    if (vel_bwd) {
        orc_window_pad3d(g, vel_bwd, oxw, nxw, oyw, nyw, b, P.vel);
        free(P.gz); free(P.gx); free(P.gy);
        higdon3(&P);
    }
    orc_f_compute_taper(tx, nxw);
    for (int k = 0; k < nyw; k++) ty[k] = 1.f;
    if (nyw > 1) orc_f_compute_taper(ty, nyw);
    counter = 0;
    for (int it = s->nt; it >= 1; it--) {
        if (it % g->snapi == 0) {   /* Imaging_condition_sourceonly_from_memory */
            counter++;
            const int index = ntsnap - counter + 1;
            const float *S = snaps + (size_t)(index - 1) * nwin;
            if (capture) {
                float *c = capture + (size_t)(counter - 1) * nwin;
                for (int k = 1; k <= nyw; k++)
                    for (int j = 1; j <= nxw; j++)
                        for (int i = 1; i <= g->nz; i++)
                            c[(i - 1) + (size_t)(j - 1) * g->nz + (size_t)(k - 1) * g->nz * nxw] = WIN(buf[3], i, j, k);
            }
            for (int k = 1; k <= nyw; k++)
                for (int j = 1; j <= nxw; j++) {
                    const float t = tx[j - 1] * ty[k - 1];
                    for (int i = 1; i <= g->nz; i++) {
                        const size_t q = (i - 1) + (size_t)(j - 1) * g->nz + (size_t)(k - 1) * g->nz * nxw;
                        img[q] = img[q] + WIN(buf[3], i, j, k) * S[q] * t;
                        ill[q] = ill[q] + S[q] * S[q] * t;
                    }
                }
        }
        step3(&P, buf, rec, s->ntr, traces_in, s->nt, it, h);
    }
#undef WIN
    /* mod_copy_image: a zeroed model-size sum of the window, then added */
    const size_t N = (size_t)g->nz * g->nx * g->ny;
    for (int pass = 0; pass < 2; pass++) {
        float *dst = pass == 0 ? image : illum;
        const float *src_w = pass == 0 ? img : ill;
        if (!dst) continue;
        float *tmp = (float *)calloc(N, sizeof(float));
        for (int k = 1; k <= nyw; k++) {
            const float cy = (float)(k - 1) * g->dy + oyw;
            const int iy = nint3((cy - g->oy) / g->dy) + 1;
            if (iy < 1 || iy > g->ny) continue;
            for (int j = 1; j <= nxw; j++) {
                const float cx = (float)(j - 1) * g->dx + oxw;
                const int ix = nint3((cx - g->ox) / g->dx) + 1;
                if (ix < 1 || ix > g->nx) continue;
                float *d = tmp + (size_t)(ix - 1) * g->nz + (size_t)(iy - 1) * g->nz * g->nx;
                const float *w = src_w + (size_t)(j - 1) * g->nz + (size_t)(k - 1) * g->nz * nxw;
                for (int i = 0; i < g->nz; i++) d[i] = d[i] + w[i];
            }
        }
        for (size_t q = 0; q < N; q++) dst[q] = dst[q] + tmp[q];
        free(tmp);
    }
    if (win_image) memcpy(win_image, img, nwin * sizeof(float));
    if (win_illum) memcpy(win_illum, ill, nwin * sizeof(float));
    if (snaps_out) memcpy(snaps_out, snaps, nwin * (size_t)ntsnap * sizeof(float));
    for (int q = 0; q < 5; q++) free(buf[q]);
    free(snaps); free(img); free(ill); free(tx); free(ty); free(wav); free(rec);
    free3(&P);
    return 0;
}

double orc_f_fct_compute(int ntype, const float *rd, int nd, float *thresh);
void orc_f_gdt_compute(int ntype, float *rd, int nd, float t);

/* progs/afwi/3DAFWI.f90's gradient, AFWI_to_memory (to_disk_to_memory_AFWI.f90:
 * 229-415) with genpar%twoD false, summed over nshots shots: per shot forward
 * modeling (Injection_sinc; readsou's half-step shift with a density) with
 * Extraction_array_sinc_afwi (gain 1, Extraction.f90:137-162 -> :276-337)
 * storing snapshots; Compute_OF_RES_3D (Objective_Function_Residual_Adjsource.
 * f90:42-67; one THRESH per shot, traces in order; the misfit starts at 0);
 * reverse time with the 3D branch's forward stencil (or the density
 * operator), Injection_sinc of the residual and
 * Imaging_condition_AFWI_from_memory (Imaging.f90:371-422), or with a density
 * Imaging_condition_AFWI_RHOVP_from_memory (:424-512, dk = 1 unless nyw = 1);
 * each window added into the model (mod_copy_gradient_to_disk,
 * ExtractPadModel.f90:592-660) and, with a density and vprho 1, the Vp/Ip
 * change of variables on the model's vel and rho.  dobs, mute: per shot
 * (nt, ntr) time fastest, concatenated (mute may be NULL); grad: (nz,nx,ny)
 * for vp, then for rho (or Ip) with a density. */
int orc_f_afwi3d(const sfd_grid3d *g, const float *vel_full, const float *rho_full, int nshots,
                 const sfd_shot3d *shots, const float *dobs, const float *mute, int norm, float norm_thresh, int vprho,
                 double *fout, float *grad, float *illum)
{
    if (g->snapi < 1) return -1;
    const size_t N = (size_t)g->nz * g->nx * g->ny;
    const int np = rho_full ? 2 : 1;
    for (size_t q = 0; q < N * np; q++) grad[q] = 0;
    for (size_t q = 0; q < N; q++) illum[q] = 0;
    double f = 0;
    size_t off = 0;
    for (int is = 0; is < nshots; is++) {
        const sfd_shot3d *s = &shots[is];
        P3 P;
        float oxw, oyw;
        int nxw, nyw;
        if (setup3(&P, g, vel_full, rho_full, s, &oxw, &nxw, &oyw, &nyw)) return -2;
        if (nxw < 2 || g->nz < 2) return -1;   /* the one-sided differences need two cells */
        const B3 *b = &P.b;
        const size_t n = (size_t)b->NZ * b->NX * b->NY, nwin = (size_t)g->nz * nxw * nyw;
        const int ntsnap = s->nt / g->snapi, h = (int)(g->lsinc * 0.5);
        float *buf[5], *snaps = (float *)calloc(nwin * (ntsnap > 0 ? ntsnap : 1), sizeof(float));
        float *img = (float *)calloc(nwin, sizeof(float)), *img2 = (float *)calloc(nwin, sizeof(float));
        float *ill = (float *)calloc(nwin, sizeof(float));
        float *tx = (float *)malloc(sizeof(float) * nxw), *ty = (float *)malloc(sizeof(float) * nyw);
        float *dmod = (float *)calloc((size_t)s->nt * s->ntr, sizeof(float));
        for (int q = 0; q < 5; q++) buf[q] = (float *)calloc(n, sizeof(float));
        if (!buf[4] || !snaps || !img || !img2 || !ill || !tx || !ty || !dmod) return -2;
        float *wav = (float *)malloc(sizeof(float) * (size_t)s->nt);
        memcpy(wav, s->wavelet, sizeof(float) * (size_t)s->nt);
        if (rho_full) orc_f_half_step(wav, s->nt, g->lsinc);
        Pt3 src, *rec = (Pt3 *)malloc(sizeof(Pt3) * (size_t)s->ntr);
        points3(g, oxw, oyw, 1, &s->sz, &s->sx, &s->sy, &src);
        points3(g, oxw, oyw, s->ntr, s->rz, s->rx, s->ry, rec);
#define WIN(u, i, j, k) A3(&P, u, i, j, k)
#define WQ(i, j, k) ((size_t)((i) - 1) + (size_t)((j) - 1) * g->nz + (size_t)((k) - 1) * g->nz * nxw)
        int counter = 0;
        for (int it = 1; it <= s->nt; it++) {
            for (int r = 0; r < s->ntr; r++) {   /* Extraction_array_sinc_afwi */
                const int *ic = rec[r].ic;
                float *tr = dmod + (size_t)r * s->nt + (it - 1), acc = *tr;
                for (int k = -h; k <= h; k++)
                    for (int j = -h; j <= h; j++)
                        for (int i = -h; i <= h; i++)
                            acc = acc + 1.f * A3(&P, buf[3], i + ic[0], j + ic[1], k + ic[2]) * rec[r].wz[h + i] *
                                            rec[r].wx[h + j] * rec[r].wx[h + k];
                *tr = acc;
            }
            if (it % g->snapi == 0) {
                float *S = snaps + (size_t)(counter++) * nwin;
                for (int k = 1; k <= nyw; k++)
                    for (int j = 1; j <= nxw; j++)
                        for (int i = 1; i <= g->nz; i++) S[WQ(i, j, k)] = WIN(buf[3], i, j, k);
            }
            step3(&P, buf, &src, 1, wav, s->nt, it, h);
        }
        /* Compute_OF_RES_3D */
        float thresh = norm_thresh;
        for (int r = 0; r < s->ntr; r++) {
            float *d = dmod + (size_t)r * s->nt;
            const float *ob = dobs + off + (size_t)r * s->nt;
            const float *mu = mute ? mute + off + (size_t)r * s->nt : NULL;
            for (int it = 0; it < s->nt; it++) d[it] = (mu ? mu[it] : 1.f) * (ob[it] - d[it]);
            f = f + orc_f_fct_compute(norm, d, s->nt, &thresh);
            for (int it = 0; it < s->nt; it++) d[it] = (mu ? mu[it] : 1.f) * d[it];
            orc_f_gdt_compute(norm, d, s->nt, thresh);
        }
        /* backward */
        for (int q = 0; q < 5; q++) memset(buf[q], 0, n * sizeof(float));
        if (P.sz) {
            memset(P.sz, 0, n * sizeof(float));
            memset(P.sx, 0, n * sizeof(float));
            memset(P.sy, 0, n * sizeof(float));
        }
        orc_f_compute_taper(tx, nxw);
        for (int k = 0; k < nyw; k++) ty[k] = 1.f;
        if (nyw > 1) orc_f_compute_taper(ty, nyw);
        const float dt2 = (g->dt * (float)g->snapi) * (g->dt * (float)g->snapi), dti2 = 1.f / dt2;
        const float dzi = 1.f / g->dz, dxi = 1.f / g->dx, dyi = 1.f / g->dy;
        const int dk = nyw == 1 ? 0 : 1;
        counter = 0;
        for (int it = s->nt; it >= 1; it--) {
            if (it % g->snapi == 0) {
                counter++;
                const int index = ntsnap - counter + 1;
                const int im1 = index - 1 > 1 ? index - 1 : 1, ip1 = index + 1 < ntsnap ? index + 1 : ntsnap;
                const float *S = snaps + (size_t)(index - 1) * nwin, *Sm = snaps + (size_t)(im1 - 1) * nwin,
                            *Sp = snaps + (size_t)(ip1 - 1) * nwin;
                for (int k = 1; k <= nyw; k++)
                    for (int j = 1; j <= nxw; j++) {
                        const float t = tx[j - 1] * ty[k - 1];
                        for (int i = 1; i <= g->nz; i++) {
                            const size_t q = WQ(i, j, k);
                            const float u = WIN(buf[3], i, j, k), v = WIN(P.vel, i, j, k);
                            if (!rho_full) {
                                const float tmp = (Sm[q] - 2.f * S[q] + Sp[q]) / (dt2 * (v * v * v));
                                img[q] = img[q] + u * tmp * t;
                            } else {
                                const int ip = i + 1 < g->nz ? i + 1 : g->nz, jp = j + 1 < nxw ? j + 1 : nxw;
                                const int kp = k + dk < nyw ? k + dk : nyw;
                                const int im = ip - 1, jm = jp - 1, km = kp - dk;
                                float tmp = dzi * (WIN(buf[3], ip, jm, km) - WIN(buf[3], im, jm, km)) * dzi *
                                            (S[WQ(ip, jm, km)] - S[WQ(im, jm, km)]);
                                tmp = tmp + dxi * (WIN(buf[3], im, jp, km) - WIN(buf[3], im, jm, km)) * dxi *
                                                (S[WQ(im, jp, km)] - S[WQ(im, jm, km)]);
                                tmp = tmp + dyi * (WIN(buf[3], im, jm, kp) - WIN(buf[3], im, jm, km)) * dyi *
                                                (S[WQ(im, jm, kp)] - S[WQ(im, jm, km)]);
                                const float tmp1 = u * dti2 * (Sm[q] - 2.f * S[q] + Sp[q]);
                                const float r = WIN(P.rho, i, j, k);
                                img[q] = img[q] + 2.f * tmp1 * t / (v * v * v * r);
                                img2[q] = img2[q] + tmp * t / (r * r) + tmp1 * t / (r * r * (v * v));
                            }
                            ill[q] = ill[q] + S[q] * S[q] * t;
                        }
                    }
            }
            step3(&P, buf, rec, s->ntr, dmod, s->nt, it, h);
        }
#undef WQ
#undef WIN
        /* mod_copy_gradient_to_disk: the window added at the nearest model columns and lines */
        for (int pass = 0; pass < 3; pass++) {
            float *dst = pass == 0 ? grad : pass == 1 ? grad + N : illum;
            const float *w = pass == 0 ? img : pass == 1 ? img2 : ill;
            if (pass == 1 && np == 1) continue;
            for (int k = 1; k <= nyw; k++) {
                const int iy = nint3(((float)(k - 1) * g->dy + oyw - g->oy) / g->dy) + 1;
                if (iy < 1 || iy > g->ny) continue;
                for (int j = 1; j <= nxw; j++) {
                    const int ix = nint3(((float)(j - 1) * g->dx + oxw - g->ox) / g->dx) + 1;
                    if (ix < 1 || ix > g->nx) continue;
                    float *d = dst + (size_t)(ix - 1) * g->nz + (size_t)(iy - 1) * g->nz * g->nx;
                    const float *ww = w + (size_t)(j - 1) * g->nz + (size_t)(k - 1) * g->nz * nxw;
                    for (int i = 0; i < g->nz; i++) d[i] = d[i] + ww[i];
                }
            }
        }
        for (int q = 0; q < 5; q++) free(buf[q]);
        free(snaps); free(img); free(img2); free(ill); free(tx); free(ty); free(dmod); free(wav); free(rec);
        free3(&P);
        off += (size_t)s->nt * s->ntr;
    }
    if (np == 2 && vprho == 1)   /* ExtractPadModel.f90:643-646 on the model's vel and rho */
        for (size_t q = 0; q < N; q++) {
            const float v = vel_full[q], r = rho_full[q];
            grad[q] = grad[q] - (r / v) * grad[N + q];
            grad[N + q] = (1.f / v) * grad[N + q];
        }
    *fout = f;
    return 0;
}

/* FDderivatives.f90:780-819 FD_2nd_3D_derivatives_scalar_adjoint_grid:
 * tmp = vel^2 u2 over the whole padded array, then u3 = (tx + tz) + ty of the
 * 25-point stencil of tmp on the FD region (no v^2 factor after). */
static void stencil3_adj(P3 *P, const float *u2, float *u3)
{
    const B3 *b = &P->b;
    const float *c = P->c;
    const size_t n = (size_t)b->NZ * b->NX * b->NY;
    float *tmp = (float *)malloc(sizeof(float) * n);
    for (size_t q = 0; q < n; q++) tmp[q] = P->vel[q] * P->vel[q] * u2[q];
#pragma omp parallel for schedule(static)
    for (int k = b->nmin3; k <= b->nmax3; k++)
        for (int j = b->nmin2; j <= b->nmax2; j++)
            for (int i = b->nmin1; i <= b->nmax1; i++) {
#define T(ii, jj, kk) A3(P, tmp, ii, jj, kk)
                float tz = c[0] * T(i, j, k) + c[1] * (T(i + 1, j, k) + T(i - 1, j, k)) +
                           c[2] * (T(i + 2, j, k) + T(i - 2, j, k)) + c[3] * (T(i + 3, j, k) + T(i - 3, j, k)) +
                           c[4] * (T(i + 4, j, k) + T(i - 4, j, k));
                float tx = c[5] * T(i, j, k) + c[6] * (T(i, j + 1, k) + T(i, j - 1, k)) +
                           c[7] * (T(i, j + 2, k) + T(i, j - 2, k)) + c[8] * (T(i, j + 3, k) + T(i, j - 3, k)) +
                           c[9] * (T(i, j + 4, k) + T(i, j - 4, k));
                float ty = c[10] * T(i, j, k) + c[11] * (T(i, j, k + 1) + T(i, j, k - 1)) +
                           c[12] * (T(i, j, k + 2) + T(i, j, k - 2)) + c[13] * (T(i, j, k + 3) + T(i, j, k - 3)) +
                           c[14] * (T(i, j, k + 4) + T(i, j, k - 4));
#undef T
                A3(P, u3, i, j, k) = tx + tz + ty;
            }
    free(tmp);
}

/* Injection_LSRTM_sinc (Injection.f90:239-272): the traces by the 3D sinc
 * with the weight 1/(dz dx dy) in place of v^2, in trace order. */
static void inject3_w(P3 *P, float *u, const Pt3 *pt, int ntr, const float *tr, int nt, int it, int h, float w)
{
    for (int r = 0; r < ntr; r++) {
        const int *ic = pt[r].ic;
        const float s = tr[(size_t)r * nt + (it - 1)];
        for (int k = -h; k <= h; k++)
            for (int j = -h; j <= h; j++)
                for (int i = -h; i <= h; i++)
                    A3(P, u, i + ic[0], j + ic[1], k + ic[2]) =
                        A3(P, u, i + ic[0], j + ic[1], k + ic[2]) + w * s * pt[r].wz[h + i] * pt[r].wx[h + j] *
                                                                        pt[r].wx[h + k];
    }
}

static void rotate3(float **buf)
{
    float *t = buf[0];
    buf[0] = buf[1]; buf[1] = buf[2]; buf[2] = buf[3]; buf[3] = buf[4]; buf[4] = t;
}

/* readsou with Born / LSRTM (ReadSouVelRho.f90:318-326): the source's second
 * time difference. */
static void second_diff3(float *tr, int nt, float dt)
{
    float *o = (float *)malloc(sizeof(float) * (size_t)nt);
    memcpy(o, tr, sizeof(float) * (size_t)nt);
    for (int it = 1; it <= nt; it++) tr[it - 1] = 0;
    for (int it = 2; it <= nt - 1; it++) tr[it - 1] = tr[it - 1] + (-2.f * o[it - 1] + o[it - 2] + o[it]) / (dt * dt);
    free(o);
}

/* The forward pass shared by LSRTM and Born in 3D: the second-differenced
 * source by Injection_sinc, the forward stencil, snapshots of u2 on the window
 * every snapi steps. */
static void fwd_snaps3(P3 *P, const sfd_grid3d *g, const sfd_shot3d *s, float oxw, float oyw, int nxw, int nyw,
                       float *snaps)
{
    const B3 *b = &P->b;
    const size_t n = (size_t)b->NZ * b->NX * b->NY, nwin = (size_t)g->nz * nxw * nyw;
    const int h = (int)(g->lsinc * 0.5);
    float *buf[5];
    for (int q = 0; q < 5; q++) buf[q] = (float *)calloc(n, sizeof(float));
    float *wav = (float *)malloc(sizeof(float) * (size_t)s->nt);
    memcpy(wav, s->wavelet, sizeof(float) * (size_t)s->nt);
    second_diff3(wav, s->nt, g->dt);
    Pt3 src;
    points3(g, oxw, oyw, 1, &s->sz, &s->sx, &s->sy, &src);
    int counter = 0;
    for (int it = 1; it <= s->nt; it++) {
        if (it % g->snapi == 0) {
            float *S = snaps + (size_t)(counter++) * nwin;
            for (int k = 1; k <= nyw; k++)
                for (int j = 1; j <= nxw; j++)
                    for (int i = 1; i <= g->nz; i++)
                        S[(i - 1) + (size_t)(j - 1) * g->nz + (size_t)(k - 1) * g->nz * nxw] = A3(P, buf[3], i, j, k);
        }
        step3(P, buf, &src, 1, wav, s->nt, it, h);
    }
    for (int q = 0; q < 5; q++) free(buf[q]);
    free(wav);
}

/* 3DRTM.x LSRTM=1 with genpar%twoD false, LSRTM_to_memory
 * (to_disk_to_memory_RTM.f90:494-579): the forward pass of fwd_snaps3, then
 * reverse time on vel_bwd when given with the adjoint-style 3D stencil,
 * Injection_LSRTM_sinc of the receivers' traces and
 * Imaging_condition_LSRTM_sourceonly_from_memory (Imaging.f90:698-745:
 * image += 2 t u S / v, illum += S S t); the window added into image / illum
 * as mod_copy_image.  Constant density only (LSRTM_to_memory has no density
 * branch).  capture (optional): the backward u2 at each imaging step. */
int orc_f_lsrtm3d(const sfd_grid3d *g, const float *vel_full, const float *vel_bwd, const sfd_shot3d *s,
                  const float *traces_in, float *image, float *illum, float *capture)
{
    P3 P;
    float oxw, oyw;
    int nxw, nyw;
    if (g->snapi < 1) return -1;
    if (setup3(&P, g, vel_full, NULL, s, &oxw, &nxw, &oyw, &nyw)) return -2;
    const B3 *b = &P.b;
    const size_t n = (size_t)b->NZ * b->NX * b->NY, nwin = (size_t)g->nz * nxw * nyw;
    const int ntsnap = s->nt / g->snapi, h = (int)(g->lsinc * 0.5);
    float *snaps = (float *)calloc(nwin * (ntsnap > 0 ? ntsnap : 1), sizeof(float));
    float *img = (float *)calloc(nwin, sizeof(float)), *ill = (float *)calloc(nwin, sizeof(float));
    float *tx = (float *)malloc(sizeof(float) * nxw), *ty = (float *)malloc(sizeof(float) * nyw), *buf[5];
    fwd_snaps3(&P, g, s, oxw, oyw, nxw, nyw, snaps);
    if (vel_bwd) {
        orc_window_pad3d(g, vel_bwd, oxw, nxw, oyw, nyw, b, P.vel);
        free(P.gz); free(P.gx); free(P.gy);
        higdon3(&P);
    }
    orc_f_compute_taper(tx, nxw);
    for (int k = 0; k < nyw; k++) ty[k] = 1.f;
    if (nyw > 1) orc_f_compute_taper(ty, nyw);
    Pt3 *rec = (Pt3 *)malloc(sizeof(Pt3) * (size_t)s->ntr);
    points3(g, oxw, oyw, s->ntr, s->rz, s->rx, s->ry, rec);
    const float w = 1.f / (g->dz * g->dx * g->dy);
    for (int q = 0; q < 5; q++) buf[q] = (float *)calloc(n, sizeof(float));
    int counter = 0;
    for (int it = s->nt; it >= 1; it--) {
        if (it % g->snapi == 0) {
            counter++;
            const int index = ntsnap - counter + 1;
            const float *S = snaps + (size_t)(index - 1) * nwin;
            float *cp = capture ? capture + (size_t)(counter - 1) * nwin : NULL;
            for (int k = 1; k <= nyw; k++)
                for (int j = 1; j <= nxw; j++) {
                    const float t = tx[j - 1] * ty[k - 1];
                    for (int i = 1; i <= g->nz; i++) {
                        const size_t q = (i - 1) + (size_t)(j - 1) * g->nz + (size_t)(k - 1) * g->nz * nxw;
                        const float u = A3(&P, buf[3], i, j, k);
                        if (cp) cp[q] = u;
                        img[q] = img[q] + 2.f * t * u * S[q] / A3(&P, P.vel, i, j, k);
                        ill[q] = ill[q] + S[q] * S[q] * t;
                    }
                }
        }
        stencil3_adj(&P, buf[3], buf[4]);
        inject3_w(&P, buf[4], rec, s->ntr, traces_in, s->nt, it, h, w);
        update3(&P, buf[0], buf[1], buf[2], buf[3], buf[4]);
        boundary3(&P, buf[2], buf[3], buf[4]);
        rotate3(buf);
    }
    const size_t N = (size_t)g->nz * g->nx * g->ny;
    for (int pass = 0; pass < 2; pass++) {
        float *dst = pass == 0 ? image : illum;
        const float *src_w = pass == 0 ? img : ill;
        if (!dst) continue;
        float *tmp = (float *)calloc(N, sizeof(float));
        for (int k = 1; k <= nyw; k++) {
            const int iy = nint3(((float)(k - 1) * g->dy + oyw - g->oy) / g->dy) + 1;
            if (iy < 1 || iy > g->ny) continue;
            for (int j = 1; j <= nxw; j++) {
                const int ix = nint3(((float)(j - 1) * g->dx + oxw - g->ox) / g->dx) + 1;
                if (ix < 1 || ix > g->nx) continue;
                float *d = tmp + (size_t)(ix - 1) * g->nz + (size_t)(iy - 1) * g->nz * g->nx;
                const float *ww = src_w + (size_t)(j - 1) * g->nz + (size_t)(k - 1) * g->nz * nxw;
                for (int i = 0; i < g->nz; i++) d[i] = d[i] + ww[i];
            }
        }
        for (size_t q = 0; q < N; q++) dst[q] = dst[q] + tmp[q];
        free(tmp);
    }
    for (int q = 0; q < 5; q++) free(buf[q]);
    free(snaps); free(img); free(ill); free(tx); free(ty); free(rec);
    free3(&P);
    return 0;
}

/* 3DBORNMOD.x with genpar%twoD false, BORN_to_memory
 * (to_disk_to_memory_BORN.f90:123-220): the forward pass of fwd_snaps3, then
 * a second forward pass (vel_bwd when given, mod%vel2) with no source: at every
 * snapi-th step Injection_Born (Imaging.f90:747-788) adds 2 t m S_counter / v
 * to u3 after the stencil, and Extraction_array_LSRTM_sinc (gain
 * 1/(dz dx dy)) records the receivers.  refl: (nz,nx,ny), windowed like vel
 * (readref + mod_window_pad, nearest columns and lines).  traces: (nt, ntr). */
int orc_f_born3d(const sfd_grid3d *g, const float *vel_full, const float *vel_bwd, const float *refl,
                 const sfd_shot3d *s, float *traces)
{
    P3 P;
    float oxw, oyw;
    int nxw, nyw;
    if (g->snapi < 1) return -1;
    if (setup3(&P, g, vel_full, NULL, s, &oxw, &nxw, &oyw, &nyw)) return -2;
    const B3 *b = &P.b;
    const size_t n = (size_t)b->NZ * b->NX * b->NY, nwin = (size_t)g->nz * nxw * nyw;
    const int ntsnap = s->nt / g->snapi, h = (int)(g->lsinc * 0.5);
    float *snaps = (float *)calloc(nwin * (ntsnap > 0 ? ntsnap : 1), sizeof(float));
    float *m = (float *)calloc(n, sizeof(float));
    float *tx = (float *)malloc(sizeof(float) * nxw), *ty = (float *)malloc(sizeof(float) * nyw), *buf[5];
    fwd_snaps3(&P, g, s, oxw, oyw, nxw, nyw, snaps);
    orc_window_pad3d(g, refl, oxw, nxw, oyw, nyw, b, m);
    if (vel_bwd) {
        orc_window_pad3d(g, vel_bwd, oxw, nxw, oyw, nyw, b, P.vel);
        free(P.gz); free(P.gx); free(P.gy);
        higdon3(&P);
    }
    orc_f_compute_taper(tx, nxw);
    for (int k = 0; k < nyw; k++) ty[k] = 1.f;
    if (nyw > 1) orc_f_compute_taper(ty, nyw);
    Pt3 *rec = (Pt3 *)malloc(sizeof(Pt3) * (size_t)s->ntr);
    points3(g, oxw, oyw, s->ntr, s->rz, s->rx, s->ry, rec);
    const float gain = 1.f / (g->dz * g->dx * g->dy);
    for (int q = 0; q < 5; q++) buf[q] = (float *)calloc(n, sizeof(float));
    memset(traces, 0, sizeof(float) * (size_t)s->nt * s->ntr);
    int counter = 0;
    for (int it = 1; it <= s->nt; it++) {
        for (int r = 0; r < s->ntr; r++) {   /* Extraction_array_LSRTM_sinc */
            const int *ic = rec[r].ic;
            float *tr = traces + (size_t)r * s->nt + (it - 1), acc = *tr;
            for (int k = -h; k <= h; k++)
                for (int j = -h; j <= h; j++)
                    for (int i = -h; i <= h; i++)
                        acc = acc + gain * A3(&P, buf[3], i + ic[0], j + ic[1], k + ic[2]) * rec[r].wz[h + i] *
                                        rec[r].wx[h + j] * rec[r].wx[h + k];
            *tr = acc;
        }
        stencil3(&P, buf[3], buf[4]);
        if (it % g->snapi == 0) {   /* Injection_Born */
            const float *S = snaps + (size_t)(counter++) * nwin;
            for (int k = 1; k <= nyw; k++)
                for (int j = 1; j <= nxw; j++) {
                    const float t = tx[j - 1] * ty[k - 1];
                    for (int i = 1; i <= g->nz; i++) {
                        const size_t q = (i - 1) + (size_t)(j - 1) * g->nz + (size_t)(k - 1) * g->nz * nxw;
                        A3(&P, buf[4], i, j, k) =
                            A3(&P, buf[4], i, j, k) + 2.f * t * A3(&P, m, i, j, k) * S[q] / A3(&P, P.vel, i, j, k);
                    }
                }
        }
        update3(&P, buf[0], buf[1], buf[2], buf[3], buf[4]);
        boundary3(&P, buf[2], buf[3], buf[4]);
        rotate3(buf);
    }
    for (int q = 0; q < 5; q++) free(buf[q]);
    free(snaps); free(m); free(tx); free(ty); free(rec);
    free3(&P);
    return 0;
}
