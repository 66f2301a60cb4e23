/*
 * seisoracle_impl.h — body of the CPU oracle, compiled twice by
 * seisoracle.c: once with REAL=float (the reference's own precision) and
 * once with REAL=double (the FP64 oracle that image/gradient errors are
 * quoted against).
 *
 * TEST INFRASTRUCTURE ONLY.  This is a line-by-line restatement of the
 * reference's 2D constant-density acoustic path in its own pass order.  It is
 * the checker for the HIP engine; nothing in the product links or calls it.
 *
 * Every routine cites the reference file:line it restates.  Arithmetic is
 * written in the reference's left-to-right evaluation order so that the FP32
 * build agrees with the flang-compiled reference kernels to rounding.
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(CAT(orc_, SFX), CAT(_, name))

/* ------------------------------------------------------------------ */
/* Fortran nint(): nearest integer, halves away from zero.            */
static int CAT(nint_, SFX)(REAL x) { return (int)ROUND(x); }
#define NINT CAT(nint_, SFX)

/* ModelSpaceTypes.f90:330-337 find_i_d */
static void CAT(find_i_d_, SFX)(int *ic, REAL *dc, REAL c, REAL o, REAL d)
{
    *ic = NINT((c - o) / d) + 1;
    *dc = c - (o + (REAL)(*ic - 1) * d);
}

/* Interpolate.f90:15-42 mksinc */
void FN(mksinc)(REAL *sinc, int lsinc, REAL shift)
{
    const REAL pi = (REAL)3.141592654;
    REAL fsum = 0, fact, pos;
    int n;
    for (n = 1; n <= lsinc; n++) {
        if (n <= (int)((REAL)(lsinc / 2 + 1) + shift))
            fact = ((REAL)(n - 1) - shift) / (REAL)(lsinc / 2);
        else
            fact = ((REAL)(lsinc - n) + shift) / (REAL)(lsinc / 2);
        if (fact <= 0) fact = 0;
        pos = pi * ((REAL)(n - (lsinc / 2 + 1)) - shift);
        if (FABS(pos) > (REAL)0.005)
            sinc[n - 1] = fact * SIN(pos) / pos;
        else
            sinc[n - 1] = fact * ((REAL)1 - pos * pos / (REAL)6);
        fsum = fsum + sinc[n - 1];
    }
    for (n = 1; n <= lsinc; n++) sinc[n - 1] = sinc[n - 1] / fsum;
}

/* Boundary.f90:66-87 Higdon_coeff */
void FN(higdon_coeff)(REAL nu, REAL beta1, REAL beta2, REAL weight, REAL *c)
{
    REAL fdum, qx, qt, qxt, rx, rt, rxt;
    fdum = (beta1 + nu) * ((REAL)1 - weight);
    qx = (weight * (beta1 + nu) - nu) / fdum;
    qt = (weight * (beta1 + nu) - beta1) / fdum;
    qxt = weight / (weight - (REAL)1);
    fdum = (beta2 + nu) * ((REAL)1 - weight);
    rx = (weight * (beta2 + nu) - nu) / fdum;
    rt = (weight * (beta2 + nu) - beta2) / fdum;
    rxt = weight / (weight - (REAL)1);
    c[0] = qx + rx;
    c[1] = qx * rx;
    c[2] = qt + rt;
    c[3] = qx * rt + qt * rx + qxt + rxt;
    c[4] = qx * rxt + rx * qxt;
    c[5] = qt * rt;
    c[6] = qt * rxt + rt * qxt;
    c[7] = qxt * rxt;
}

/* Taper.f90:30-44 cos_sin_taper, applied as compute_taper does (:14-28) */
void FN(compute_taper)(REAL *t, int n)
{
    const REAL pi = (REAL)3.14159265359;
    int nt = (n / 2 < 10) ? n / 2 : 10;
    int m = nt < n ? nt : n, i, b;
    for (i = 0; i < n; i++) t[i] = 1;
    for (i = 1; i <= m; i++) {
        REAL s = SIN(pi * (REAL)(i - 1) / (REAL)(2 * (m - 1)));
        t[i - 1] = s * s;
    }
    b = (n - nt) > 1 ? (n - nt) : 1;
    for (i = b; i <= n; i++) {
        REAL c = COS(pi * (REAL)(i - b) / (REAL)(2 * nt));
        t[i - 1] = c * c;
    }
}

/* ------------------------------------------------------------------ */
typedef struct {
    REAL coord[2];   /* z, x */
    REAL dcoord[2];
    int icoord[2];
    REAL *trace;     /* nt samples, trace(it) at trace[it-1] */
} CAT(TR_, SFX);
#define TR CAT(TR_, SFX)

typedef struct {
    ORC_BOUNDS b;
    int nz, nxw;             /* window interior                         */
    REAL oz, oxw, dz, dx, dt;
    int lsinc, snapi, ntsnap;
    REAL *vel;               /* padded vel, NZ*NX                        */
    REAL c[10];              /* c0z..c4z, c0x..c4x (FDtypes.f90:367-387) */
    REAL *gz, *gx;           /* Higdon tables                            */
    int stencil;             /* 0 forward v^2 lap(u), 1 adjoint-style lap(v^2 u) (with_rho: the
                                variable-density operator either way, as the drivers call it) */
    int inject;              /* 0 none, 1 sinc v^2, 2 simple v^2, 3 sinc 1/(dz dx) */
    int nsou; TR *sou;
    int extract;             /* -1 none, else SFD_EXTRACT_*              */
    int ndat; TR *dat;
    REAL *wave; int wcounter;          /* snapshots nz*nxw*ntsnap, NULL = none */
    int imaging;             /* 0 none, 1 RTM, 2 FWI, 3 LSRTM, 4 Born injection, 5 FWI for vp and rho */
    REAL *image, *illum;     /* nz*nxw window accumulators               */
    REAL *image2;            /* imaging 5: the rho gradient (imagesmall_nparam(:,:,:,2)) */
    const REAL *born_m;      /* nz*nxw reflectivity (Born)               */
    const REAL *snaps;       /* forward snapshots read by the imaging   */
    REAL *taperx;
    REAL dt2snap;            /* genpar%dt2 = (dt*snapi)^2                */
    int icounter;
    int tmin, tmax, tstep;
    REAL *capture; int ncapture;  /* optional: u2 at imaging steps (checks) */
    int rec_type, shot_type, surf_type;
    int *ielev; REAL *delev;      /* per FD column nmin2..nmax2 (surf_type != 0) */
    int with_rho;                 /* variable density (genpar%withRho) */
    REAL *rho, *delp;             /* padded rho and 1/rho2 (NZ*NX) */
    REAL *s2dzz, *s2dxx;          /* staggered gradients (NZ*NX), persistent over the steps */
} CAT(PROP_, SFX);
#define PROP CAT(PROP_, SFX)

#define AT(P, u, i, j) (u)[ORC_IDX(&(P)->b, i, j)]

/* FDtypes.f90:50-80 scaled coefficients, FDcoefs.f90:25-36 values; with
 * variable density FDcoefs.f90:12-23 (staggered first derivative) scaled by
 * 1/delta (coefpower 1, ReadParam.f90:49-53) */
static void CAT(init_coefs_, SFX)(PROP *P)
{
    if (P->with_rho) {
        REAL r[5] = {(REAL)0., (REAL)1225. / (REAL)1024., -((REAL)1225. / ((REAL)1024. * (REAL)15.)),
                     (REAL)1225. / ((REAL)1024. * (REAL)125.), -((REAL)1225. / ((REAL)1024. * (REAL)1715.))};
        REAL dzi = (REAL)1. / P->dz, dxi = (REAL)1. / P->dx;
        for (int k = 0; k < 5; k++) {
            P->c[k] = r[k] * dzi;
            P->c[5 + k] = r[k] * dxi;
        }
        return;
    }
    REAL c0 = (REAL)-205. / (REAL)72., c1 = (REAL)8. / (REAL)5, c2 = (REAL)-1. / (REAL)5.,
         c3 = (REAL)8. / (REAL)315., c4 = (REAL)-1. / (REAL)560.;
    REAL dzi = (REAL)1. / P->dz, dxi = (REAL)1. / P->dx;
    REAL dz2 = dzi * dzi, dx2 = dxi * dxi;   /* dzi**coefpower, coefpower=2 */
    P->c[0] = c0 * dz2; P->c[1] = c1 * dz2; P->c[2] = c2 * dz2; P->c[3] = c3 * dz2; P->c[4] = c4 * dz2;
    P->c[5] = c0 * dx2; P->c[6] = c1 * dx2; P->c[7] = c2 * dx2; P->c[8] = c3 * dx2; P->c[9] = c4 * dx2;
}

/* Boundary.f90:17-63 Higdon (2D part) */
static void CAT(higdon_tables_, SFX)(PROP *P)
{
    const ORC_BOUNDS *b = &P->b;
    REAL dxi = (REAL)1. / P->dx, dzi = (REAL)1. / P->dz, nu;
    int j, i;
    P->gz = (REAL *)calloc((size_t)16 * (b->nmax2 - b->nmin2 + 1), sizeof(REAL));
    P->gx = (REAL *)calloc((size_t)16 * (b->nmax1 - b->nmin1 + 1), sizeof(REAL));
    for (j = b->nmin2; j <= b->nmax2; j++) {
        REAL *g = P->gz + 16 * (j - b->nmin2);
        nu = AT(P, P->vel, b->nmin1, j) * P->dt * dzi;
        FN(higdon_coeff)(nu, (REAL)1.0, (REAL)0.7, (REAL)0.5, g);
        nu = AT(P, P->vel, b->nmax1, j) * P->dt * dzi;
        FN(higdon_coeff)(nu, (REAL)1.0, (REAL)0.7, (REAL)0.5, g + 8);
    }
    for (i = b->nmin1; i <= b->nmax1; i++) {
        REAL *g = P->gx + 16 * (i - b->nmin1);
        nu = AT(P, P->vel, i, b->nmin2) * P->dt * dxi;
        FN(higdon_coeff)(nu, (REAL)1.0, (REAL)0.7, (REAL)0.5, g);
        nu = AT(P, P->vel, i, b->nmax2) * P->dt * dxi;
        FN(higdon_coeff)(nu, (REAL)1.0, (REAL)0.7, (REAL)0.5, g + 8);
    }
}

/* FDderivatives.f90:715-743 FD_2nd_2D_derivatives_scalar_forward_grid */
static void CAT(fd_forward_, SFX)(PROP *P, const REAL *u2, REAL *u3)
{
    const ORC_BOUNDS *b = &P->b;
    const REAL *c = P->c;
    /* columns are independent (no reduction): threads do not change results */
#pragma omp parallel for schedule(static)
    for (int j = b->nmin2; j <= b->nmax2; j++)
        for (int i = b->nmin1; i <= b->nmax1; i++) {
            REAL tz = c[0] * (AT(P, u2, i, j)) + c[1] * (AT(P, u2, i + 1, j) + AT(P, u2, i - 1, j)) +
                      c[2] * (AT(P, u2, i + 2, j) + AT(P, u2, i - 2, j)) +
                      c[3] * (AT(P, u2, i + 3, j) + AT(P, u2, i - 3, j)) +
                      c[4] * (AT(P, u2, i + 4, j) + AT(P, u2, i - 4, j));
            REAL tx = c[5] * (AT(P, u2, i, j)) + c[6] * (AT(P, u2, i, j + 1) + AT(P, u2, i, j - 1)) +
                      c[7] * (AT(P, u2, i, j + 2) + AT(P, u2, i, j - 2)) +
                      c[8] * (AT(P, u2, i, j + 3) + AT(P, u2, i, j - 3)) +
                      c[9] * (AT(P, u2, i, j + 4) + AT(P, u2, i, j - 4));
            REAL v = AT(P, P->vel, i, j);
            AT(P, u3, i, j) = v * v * (tx + tz);
        }
}

/* FDderivatives.f90:847-891 FD_2nd_2D_derivatives_scalar_adjoint_grid:
 * lap(v^2 u) with the velocity index clamped to the FD region. */
static void CAT(fd_adjoint_, SFX)(PROP *P, const REAL *u2, REAL *u3)
{
    const ORC_BOUNDS *b = &P->b;
    const REAL *c = P->c;
#define V2(ii, jj) (AT(P, P->vel, ii, jj) * AT(P, P->vel, ii, jj))
#pragma omp parallel for schedule(static)
    for (int j = b->nmin2; j <= b->nmax2; j++) {
        int jp[5], jm[5], k;
        for (k = 1; k <= 4; k++) {
            jp[k] = (j + k) < b->nmax2 ? (j + k) : b->nmax2;
            jm[k] = (j - k) > b->nmin2 ? (j - k) : b->nmin2;
        }
        for (int i = b->nmin1; i <= b->nmax1; i++) {
            int ip[5], im[5];
            for (k = 1; k <= 4; k++) {
                ip[k] = (i + k) < b->nmax1 ? (i + k) : b->nmax1;
                im[k] = (i - k) > b->nmin1 ? (i - k) : b->nmin1;
            }
            REAL tz = c[0] * (V2(i, j) * AT(P, u2, i, j)) +
                      c[1] * (V2(ip[1], j) * AT(P, u2, i + 1, j) + V2(im[1], j) * AT(P, u2, i - 1, j)) +
                      c[2] * (V2(ip[2], j) * AT(P, u2, i + 2, j) + V2(im[2], j) * AT(P, u2, i - 2, j)) +
                      c[3] * (V2(ip[3], j) * AT(P, u2, i + 3, j) + V2(im[3], j) * AT(P, u2, i - 3, j)) +
                      c[4] * (V2(ip[4], j) * AT(P, u2, i + 4, j) + V2(im[4], j) * AT(P, u2, i - 4, j));
            REAL tx = c[5] * (V2(i, j) * AT(P, u2, i, j)) +
                      c[6] * (V2(i, jp[1]) * AT(P, u2, i, j + 1) + V2(i, jm[1]) * AT(P, u2, i, j - 1)) +
                      c[7] * (V2(i, jp[2]) * AT(P, u2, i, j + 2) + V2(i, jm[2]) * AT(P, u2, i, j - 2)) +
                      c[8] * (V2(i, jp[3]) * AT(P, u2, i, j + 3) + V2(i, jm[3]) * AT(P, u2, i, j - 3)) +
                      c[9] * (V2(i, jp[4]) * AT(P, u2, i, j + 4) + V2(i, jm[4]) * AT(P, u2, i, j - 4));
            AT(P, u3, i, j) = tx + tz;
        }
    }
#undef V2
}

/* FDderivatives.f90:588-657 FD_2D_derivatives_acoustic_forward_grid: staggered
 * gradients times 1/rho2 on the FD region, first-order gradients times the
 * boundary's 1/rho2 in the operator padding (3 rows / columns; the corners
 * and the outermost row / column are never computed and keep their initial
 * zero), then the staggered divergence times v^2 rho. */
static void CAT(fd_rho_, SFX)(PROP *P, const REAL *u2, REAL *u3)
{
    const ORC_BOUNDS *b = &P->b;
    const REAL *c = P->c;
    REAL *sz = P->s2dzz, *sx = P->s2dxx;
    const REAL dzi = (REAL)1. / P->dz, dxi = (REAL)1. / P->dx;
#define D(i, j) AT(P, P->delp, i, j)
#define U(i, j) AT(P, u2, i, j)
#pragma omp parallel for schedule(static)
    for (int j = b->nmin2; j <= b->nmax2; j++)
        for (int i = b->nmin1; i <= b->nmax1; i++) {
            AT(P, sz, i, j) = (c[1] * (U(i + 1, j) - U(i, j)) + c[2] * (U(i + 2, j) - U(i - 1, j)) +
                               c[3] * (U(i + 3, j) - U(i - 2, j)) + c[4] * (U(i + 4, j) - U(i - 3, j))) * D(i, j);
            AT(P, sx, i, j) = (c[6] * (U(i, j + 1) - U(i, j)) + c[7] * (U(i, j + 2) - U(i, j - 1)) +
                               c[8] * (U(i, j + 3) - U(i, j - 2)) + c[9] * (U(i, j + 4) - U(i, j - 3))) * D(i, j);
        }
    for (int j = b->nmin2; j <= b->nmax2; j++) {
        for (int i = b->nmin1 - 1; i >= b->nmin1 - 3; i--) {
            AT(P, sz, i, j) = dzi * (U(i + 1, j) - U(i, j)) * D(b->nmin1, j);
            AT(P, sx, i, j) = dxi * (U(i, j + 1) - U(i, j)) * D(b->nmin1, j);
        }
        for (int i = b->nmax1 + 1; i <= b->nmax1 + 3; i++) {
            AT(P, sz, i, j) = dzi * (U(i + 1, j) - U(i, j)) * D(b->nmax1, j);
            AT(P, sx, i, j) = dxi * (U(i, j + 1) - U(i, j)) * D(b->nmax1, j);
        }
    }
    for (int j = b->nmin2 - 1; j >= b->nmin2 - 3; j--)
        for (int i = b->nmin1; i <= b->nmax1; i++) {
            AT(P, sz, i, j) = dzi * (U(i + 1, j) - U(i, j)) * D(i, b->nmin2);
            AT(P, sx, i, j) = dxi * (U(i, j + 1) - U(i, j)) * D(i, b->nmin2);
        }
    for (int j = b->nmax2 + 1; j <= b->nmax2 + 3; j++)
        for (int i = b->nmin1; i <= b->nmax1; i++) {
            AT(P, sz, i, j) = dzi * (U(i + 1, j) - U(i, j)) * D(i, b->nmax2);
            AT(P, sx, i, j) = dxi * (U(i, j + 1) - U(i, j)) * D(i, b->nmax2);
        }
#undef U
#undef D
#pragma omp parallel for schedule(static)
    for (int j = b->nmin2; j <= b->nmax2; j++)
        for (int i = b->nmin1; i <= b->nmax1; i++) {
            REAL tz = c[1] * (AT(P, sz, i, j) - AT(P, sz, i - 1, j)) + c[2] * (AT(P, sz, i + 1, j) - AT(P, sz, i - 2, j)) +
                      c[3] * (AT(P, sz, i + 2, j) - AT(P, sz, i - 3, j)) + c[4] * (AT(P, sz, i + 3, j) - AT(P, sz, i - 4, j));
            REAL tx = c[6] * (AT(P, sx, i, j) - AT(P, sx, i, j - 1)) + c[7] * (AT(P, sx, i, j + 1) - AT(P, sx, i, j - 2)) +
                      c[8] * (AT(P, sx, i, j + 2) - AT(P, sx, i, j - 3)) + c[9] * (AT(P, sx, i, j + 3) - AT(P, sx, i, j - 4));
            REAL v = AT(P, P->vel, i, j);
            AT(P, u3, i, j) = v * v * AT(P, P->rho, i, j) * (tx + tz);
        }
}

/* Interpolate.f90:44-130 Interpolate, 2D: rho2 = rho sinc-interpolated half a
 * cell along z, then along x (9-point sinc at shift 0.5, ends extended by
 * their edge values); delp = 1/rho2 (allocate_sxx2d_szz2d_delp2d,
 * FDderivatives.f90:574-582). */
static void CAT(rho_half_, SFX)(PROP *P)
{
    const ORC_BOUNDS *b = &P->b;
    const int L = 9, h = L / 2;
    REAL sinc[9];
    FN(mksinc)(sinc, L, (REAL)0.5);
    const int i0 = b->nmin1 - 4, i1 = b->nmax1 + 4, j0 = b->nmin2 - 4, j1 = b->nmax2 + 4;
    const int nmax = (i1 - i0 > j1 - j0 ? i1 - i0 : j1 - j0) + 1 + 2 * L;
    REAL *tr = (REAL *)malloc(sizeof(REAL) * (size_t)nmax), *r2 = P->delp;
    for (int j = j0; j <= j1; j++) {
        REAL *t = tr + L - i0;   /* t[i] for i in i0-L .. i1+L */
        for (int i = i0; i <= i1; i++) t[i] = AT(P, P->rho, i, j);
        for (int i = i0 - L; i <= i0 - 1; i++) t[i] = t[i0];
        for (int i = i1 + 1; i <= i1 + L; i++) t[i] = t[i1];
        for (int i = i0; i <= i1; i++) {
            REAL f = 0;
            for (int l = -h; l <= h; l++) f = f + sinc[h + l] * t[i + l];
            AT(P, r2, i, j) = f;
        }
    }
    for (int i = i0; i <= i1; i++) {
        REAL *t = tr + L - j0;
        for (int j = j0; j <= j1; j++) t[j] = AT(P, r2, i, j);
        for (int j = j0 - L; j <= j0 - 1; j++) t[j] = t[j0];
        for (int j = j1 + 1; j <= j1 + L; j++) t[j] = t[j1];
        for (int j = j0; j <= j1; j++) {
            REAL f = 0;
            for (int l = -h; l <= h; l++) f = f + sinc[h + l] * t[j + l];
            AT(P, r2, i, j) = f;
        }
    }
    free(tr);
    for (size_t q = 0; q < (size_t)b->NZ * b->NX; q++) r2[q] = (REAL)1. / r2[q];
}

/* readsou with withRho (ReadSouVelRho.f90:328-349): the source shifted half a
 * time step forward by the lsinc-point sinc, renormalised where the
 * interpolator runs off the trace */
static void CAT(half_step_, SFX)(REAL *tr, int nt, int lsinc)
{
    REAL sinc[64], *o = (REAL *)malloc(sizeof(REAL) * (size_t)nt);
    const int h = lsinc / 2;
    FN(mksinc)(sinc, lsinc, (REAL)0.5);
    memcpy(o, tr, sizeof(REAL) * (size_t)nt);
    for (int it = 1; it <= nt; it++) {
        REAL fscale = 1, fdum = 0;
        for (int l = -h; l <= h; l++) {
            if (it + l >= 1 && it + l <= nt) fdum = fdum + sinc[h + l] * o[it + l - 1];
            else fscale = fscale - sinc[h + l];
        }
        tr[it - 1] = fdum / fscale;
    }
    free(o);
}
void FN(half_step)(REAL *tr, int nt, int lsinc) { CAT(half_step_, SFX)(tr, nt, lsinc); }

/* FDderivatives.f90:960-982 FD_2nd_time_derivative_grid */
static void CAT(time_update_, SFX)(PROP *P, const REAL *u_1, const REAL *u0, const REAL *u1,
                                   const REAL *u2, REAL *u3)
{
    const ORC_BOUNDS *b = &P->b;
    REAL div11 = (REAL)1. / (REAL)11, dt2 = P->dt * P->dt;
#pragma omp parallel for schedule(static)
    for (int j = b->nmin2; j <= b->nmax2; j++)
        for (int i = b->nmin1; i <= b->nmax1; i++) {
            size_t q = ORC_IDX(b, i, j);
            u3[q] = ((REAL)20. * u2[q] - (REAL)6. * u1[q] - (REAL)4. * u0[q] + u_1[q] +
                     (REAL)12. * dt2 * u3[q]) * div11;
        }
}

/* Boundary.f90:366-447 Higdon_factor with (u0,u1,u2) = (grid%u1, grid%u2, grid%u3) */
static void CAT(higdon_factor_, SFX)(PROP *P, const REAL *U1, const REAL *U2, REAL *U3,
                                     int i, int j, int axis)
{
    const ORC_BOUNDS *b = &P->b;
    REAL f;
    const REAL *g;
    switch (axis) {
    case -1:
        g = P->gz + 16 * (j - b->nmin2);
        f = g[0] * AT(P, U3, i + 1, j) + g[1] * AT(P, U3, i + 2, j) + g[2] * AT(P, U2, i, j) +
            g[3] * AT(P, U2, i + 1, j) + g[4] * AT(P, U2, i + 2, j) + g[5] * AT(P, U1, i, j) +
            g[6] * AT(P, U1, i + 1, j) + g[7] * AT(P, U1, i + 2, j);
        break;
    case 1:
        g = P->gz + 16 * (j - b->nmin2) + 8;
        f = g[0] * AT(P, U3, i - 1, j) + g[1] * AT(P, U3, i - 2, j) + g[2] * AT(P, U2, i, j) +
            g[3] * AT(P, U2, i - 1, j) + g[4] * AT(P, U2, i - 2, j) + g[5] * AT(P, U1, i, j) +
            g[6] * AT(P, U1, i - 1, j) + g[7] * AT(P, U1, i - 2, j);
        break;
    case -2:
        g = P->gx + 16 * (i - b->nmin1);
        f = g[0] * AT(P, U3, i, j + 1) + g[1] * AT(P, U3, i, j + 2) + g[2] * AT(P, U2, i, j) +
            g[3] * AT(P, U2, i, j + 1) + g[4] * AT(P, U2, i, j + 2) + g[5] * AT(P, U1, i, j) +
            g[6] * AT(P, U1, i, j + 1) + g[7] * AT(P, U1, i, j + 2);
        break;
    default: /* +2 */
        g = P->gx + 16 * (i - b->nmin1) + 8;
        f = g[0] * AT(P, U3, i, j - 1) + g[1] * AT(P, U3, i, j - 2) + g[2] * AT(P, U2, i, j) +
            g[3] * AT(P, U2, i, j - 1) + g[4] * AT(P, U2, i, j - 2) + g[5] * AT(P, U1, i, j) +
            g[6] * AT(P, U1, i, j - 1) + g[7] * AT(P, U1, i, j - 2);
        break;
    }
    AT(P, U3, i, j) = -f;
}

/* Boundary.f90:89-184 Boundary_3d, 2D (nstencil = 3): with surf_type > 0
 * the top band is dropped and the side bands start at row nz0 = 1. */
static void CAT(boundary_, SFX)(PROP *P, const REAL *u1, const REAL *u2, REAL *u3)
{
    const ORC_BOUNDS *b = &P->b;
    int i, j, nz0 = P->surf_type > 0 ? 1 : b->nmin1;
    for (j = b->nmin2; j <= b->nmax2; j++) {
        if (P->surf_type <= 0)
            for (i = nz0 + 3; i >= b->nmin1 - 4; i--) CAT(higdon_factor_, SFX)(P, u1, u2, u3, i, j, -1);
        for (i = b->nmax1 - 3; i <= b->nmax1 + 4; i++) CAT(higdon_factor_, SFX)(P, u1, u2, u3, i, j, 1);
    }
    for (j = b->nmin2 + 3; j >= b->nmin2 - 4; j--)
        for (i = nz0; i <= b->nmax1; i++) CAT(higdon_factor_, SFX)(P, u1, u2, u3, i, j, -2);
    for (j = b->nmax2 - 3; j <= b->nmax2 + 4; j++)
        for (i = nz0; i <= b->nmax1; i++) CAT(higdon_factor_, SFX)(P, u1, u2, u3, i, j, 2);
}

/* Injection.f90:151-194 Injection_sinc -> :402-456 Injection_source_sinc_xz,
 * :63-104 Injection_simple_noomp -> :561-589, :239-272 Injection_LSRTM_sinc.
 * Ghosts: the sinc form tests shot_type (:432), the simple form type_inject
 * = rec_type when tmax is 1, else shot_type (:74-78); the mirror term is
 * subtracted at row -(row) right after the direct one (:447-450, :583-586). */
static void CAT(inject_, SFX)(PROP *P, REAL *u, int it)
{
    int k, i, j, l = P->lsinc;
    int minx = (int)(-l * 0.5), maxx = (int)(l * 0.5);
    int type_inject = P->tmax == 1 ? P->rec_type : P->shot_type;
    REAL sz[64], sx[64];
    for (k = 0; k < P->nsou; k++) {
        TR *s = &P->sou[k];
        REAL v2r, tr = s->trace[it - 1];
        if (P->inject == 3) {
            v2r = (REAL)1 / (P->dz * P->dx);
        } else {
            REAL v = AT(P, P->vel, s->icoord[0], s->icoord[1]);
            v2r = v * v;
            if (P->with_rho) v2r = v2r * AT(P, P->rho, s->icoord[0], s->icoord[1]);   /* Injection.f90:172-176 */
        }
        if (P->inject == 2) {
            AT(P, u, s->icoord[0], s->icoord[1]) = AT(P, u, s->icoord[0], s->icoord[1]) + v2r * tr;
            if (type_inject != 0)
                AT(P, u, -s->icoord[0], s->icoord[1]) = AT(P, u, -s->icoord[0], s->icoord[1]) - v2r * tr;
            continue;
        }
        FN(mksinc)(sz, l, s->dcoord[0] / P->dz);
        FN(mksinc)(sx, l, s->dcoord[1] / P->dx);
        for (j = minx; j <= maxx; j++)
            for (i = minx; i <= maxx; i++) {
                AT(P, u, i + s->icoord[0], j + s->icoord[1]) =
                    AT(P, u, i + s->icoord[0], j + s->icoord[1]) + v2r * tr * sz[maxx + i] * sx[maxx + j];
                if (P->shot_type != 0)
                    AT(P, u, i - s->icoord[0], j + s->icoord[1]) =
                        AT(P, u, i - s->icoord[0], j + s->icoord[1]) - v2r * tr * sz[maxx - i] * sx[maxx + j];
            }
    }
}

/* Extraction.f90:109-134 (sinc, v^2), :71-106 (nearest afwi), :217-246 (LSRTM),
 * :137-162 (sinc afwi, gain 1) -> :511-560 Extraction_1trace_sinc_xy */
static void CAT(extract_, SFX)(PROP *P, const REAL *u, int it)
{
    int k, i, j, l = P->lsinc;
    int minx = (int)(-l * 0.5), maxx = (int)(l * 0.5);
    REAL sz[64], sx[64];
    for (k = 0; k < P->ndat; k++) {
        TR *d = &P->dat[k];
        REAL g, acc;
        if (P->extract == SFD_EXTRACT_NEAREST) {
            if (P->rec_type != 0)   /* Extraction.f90:88-90 */
                d->trace[it - 1] = d->trace[it - 1] + (AT(P, u, d->icoord[0], d->icoord[1]) -
                                                       AT(P, u, -d->icoord[0], d->icoord[1]));
            else
                d->trace[it - 1] = d->trace[it - 1] + AT(P, u, d->icoord[0], d->icoord[1]);
            continue;
        }
        if (P->extract == SFD_EXTRACT_SINC_V2) {
            REAL v = AT(P, P->vel, d->icoord[0], d->icoord[1]);
            g = v * v;
        } else if (P->extract == SFD_EXTRACT_SINC_LSRTM) {
            g = (REAL)1 / (P->dz * P->dx);
        } else {
            g = 1;
        }
        FN(mksinc)(sz, l, d->dcoord[0] / P->dz);
        FN(mksinc)(sx, l, d->dcoord[1] / P->dx);
        acc = d->trace[it - 1];
        if (P->rec_type != 0) {   /* Extraction.f90:548-556 */
            for (j = minx; j <= maxx; j++)
                for (i = minx; i <= maxx; i++)
                    acc = acc + g * (AT(P, u, i + d->icoord[0], j + d->icoord[1]) -
                                     AT(P, u, -i - d->icoord[0], j + d->icoord[1])) * sz[maxx + i] * sx[maxx + j];
        } else {
            for (j = minx; j <= maxx; j++)
                for (i = minx; i <= maxx; i++)
                    acc = acc + g * AT(P, u, i + d->icoord[0], j + d->icoord[1]) * sz[maxx + i] * sx[maxx + j];
        }
        d->trace[it - 1] = acc;
    }
}

/* Extraction.f90:562-601 Extraction_wavefield: with a free surface the rows
 * above ielev_z of each column are zero */
static void CAT(extract_wave_, SFX)(PROP *P, const REAL *u, int it)
{
    if (it % P->snapi != 0) return;
    P->wcounter++;
    if (P->wcounter > P->ntsnap) return;
    REAL *w = P->wave + (size_t)(P->wcounter - 1) * P->nz * P->nxw;
#pragma omp parallel for schedule(static)
    for (int jj = 1; jj <= P->nxw; jj++) {
        const int i0 = P->surf_type != 0 ? P->ielev[jj - P->b.nmin2] : 1;
        for (int ii = 1; ii <= P->nz; ii++)
            w[(ii - 1) + (size_t)(jj - 1) * P->nz] = ii >= i0 ? AT(P, u, ii, jj) : (REAL)0;
    }
}

/* Boundary.f90:1609-1658 Boundary_set_free_surface_grid on u2 */
static void CAT(free_surface_, SFX)(PROP *P, REAL *u)
{
    const ORC_BOUNDS *b = &P->b;
    const int h = P->lsinc / 2;
    REAL sinc[64], ui[64];
    if (P->surf_type == 0) return;
    for (int j = b->nmin2; j <= b->nmax2; j++) {
        const int ie = P->ielev[j - b->nmin2];
        const REAL de = P->delev[j - b->nmin2];
        for (int i = b->nmin1; i <= ie - 1; i++) AT(P, u, i, j) = 0;
        if (de == 0) {
            for (int i = 0; i <= 2; i++) AT(P, u, ie - 2 - i, j) = -AT(P, u, ie + i, j);
        } else {
            FN(mksinc)(sinc, P->lsinc, de / P->dz);
            for (int i = -h; i <= h; i++) {
                REAL f = 0;
                for (int k = -h; k <= h; k++) f = f + AT(P, u, ie + k + i, j) * sinc[h + k];
                ui[h + i] = f;
            }
            for (int i = 2; i <= h + 1; i++)
                for (int k = -h; k <= i - 1; k++)
                    AT(P, u, ie - i + k, j) = AT(P, u, ie - i + k, j) - ui[h + i - 2] * sinc[h + k];
        }
    }
}

/* Imaging.f90:228-275 (RTM), :322-370 (FWI), :698-745 (LSRTM), :747-788 (Born) */
static void CAT(imaging_, SFX)(PROP *P, REAL *u, int it)
{
    int i, j, idx, im1, ip1;
    size_t n = (size_t)P->nz * P->nxw;
    if (it % P->snapi != 0) return;
    P->icounter++;
    if (P->imaging == 4) {
        /* Injection_Born: u3 += 2*taper*m*S(counter)/v in forward time order */
        const REAL *S = P->snaps + (size_t)(P->icounter - 1) * n;
#pragma omp parallel for schedule(static)
        for (int j = 1; j <= P->nxw; j++) {
            REAL t = P->taperx[j - 1];
            const int i0 = P->surf_type != 0 ? P->ielev[j - P->b.nmin2] : 1;   /* Imaging.f90:762-770 */
            for (int i = i0; i <= P->nz; i++) {
                size_t q = (i - 1) + (size_t)(j - 1) * P->nz;
                AT(P, u, i, j) = AT(P, u, i, j) + (REAL)2 * t * P->born_m[q] * S[q] / AT(P, P->vel, i, j);
            }
        }
        return;
    }
    idx = P->ntsnap - P->icounter + 1;
    ip1 = idx + 1 < P->ntsnap ? idx + 1 : P->ntsnap;
    im1 = idx - 1 > 1 ? idx - 1 : 1;
    if (idx < 1) return;
    const REAL *S = P->snaps + (size_t)(idx - 1) * n;
    const REAL *Sm = P->snaps + (size_t)(im1 - 1) * n;
    const REAL *Sp = P->snaps + (size_t)(ip1 - 1) * n;
    if (P->capture && P->icounter <= P->ncapture) {
        REAL *c = P->capture + (size_t)(P->icounter - 1) * n;
        for (j = 1; j <= P->nxw; j++)
            for (i = 1; i <= P->nz; i++) c[(i - 1) + (size_t)(j - 1) * P->nz] = AT(P, u, i, j);
    }
#pragma omp parallel for schedule(static)
    for (int jj = 1; jj <= P->nxw; jj++) {
        const int j = jj;
        REAL t = P->taperx[j - 1];
        const int i0 = P->surf_type != 0 ? P->ielev[j - P->b.nmin2] : 1;   /* Imaging.f90:247-256 */
        for (int i = i0; i <= P->nz; i++) {
            size_t q = (i - 1) + (size_t)(j - 1) * P->nz;
            REAL uu = AT(P, u, i, j);
            if (P->imaging == 1) {
                P->image[q] = P->image[q] + uu * S[q] * t;
            } else if (P->imaging == 5) {
                /* Imaging.f90:424-512 Imaging_condition_AFWI_RHOVP_from_memory in
                 * 2D (dk = 0): one-sided gradients of u and of the snapshot,
                 * ip = min(nz, i+1), jp = min(j+1, nxw), im = ip-1, jm = jp-1;
                 * the y term is u and S minus themselves, a zero */
                const REAL dzi = (REAL)1. / P->dz, dxi = (REAL)1. / P->dx, dti2 = (REAL)1. / P->dt2snap;
                const int ip = i + 1 < P->nz ? i + 1 : P->nz, jp = j + 1 < P->nxw ? j + 1 : P->nxw;
                const int im = ip - 1, jm = jp - 1;
#define WS(a, b) S[((a) - 1) + (size_t)((b) - 1) * P->nz]
                REAL tmp = dzi * (AT(P, u, ip, jm) - AT(P, u, im, jm)) * dzi * (WS(ip, jm) - WS(im, jm));
                tmp = tmp + dxi * (AT(P, u, im, jp) - AT(P, u, im, jm)) * dxi * (WS(im, jp) - WS(im, jm));
                tmp = tmp + (REAL)0;
#undef WS
                const REAL tmp1 = uu * dti2 * (Sm[q] - (REAL)2 * S[q] + Sp[q]);
                const REAL v = AT(P, P->vel, i, j), r = AT(P, P->rho, i, j);
                P->image[q] = P->image[q] + (REAL)2 * tmp1 * t / (v * v * v * r);
                P->image2[q] = P->image2[q] + tmp * t / (r * r) + tmp1 * t / (r * r * (v * v));
            } else if (P->imaging == 2) {
                REAL v = AT(P, P->vel, i, j);
                REAL tmp = (Sm[q] - (REAL)2 * S[q] + Sp[q]) / (P->dt2snap * (v * v * v));
                P->image[q] = P->image[q] + uu * tmp * t;
            } else { /* LSRTM */
                P->image[q] = P->image[q] + (REAL)2 * t * uu * S[q] / AT(P, P->vel, i, j);
            }
            P->illum[q] = P->illum[q] + S[q] * S[q] * t;
        }
    }
}

/* Propagator.f90:27-270 propagator_acoustic, the time loop of :168-243 */
static void CAT(propagate_, SFX)(PROP *P)
{
    const ORC_BOUNDS *b = &P->b;
    size_t n = (size_t)b->NZ * b->NX;
    REAL *buf[5], *tmp;
    int k, it;
    for (k = 0; k < 5; k++) buf[k] = (REAL *)calloc(n, sizeof(REAL));
    /* buf[0]=u_1, buf[1]=u0, buf[2]=u1, buf[3]=u2, buf[4]=u3 */
    if (P->with_rho) {   /* allocate_sxx2d_szz2d_delp2d: cells never written stay 0 */
        P->s2dzz = (REAL *)calloc(n, sizeof(REAL));
        P->s2dxx = (REAL *)calloc(n, sizeof(REAL));
    }
    CAT(init_coefs_, SFX)(P);
    CAT(higdon_tables_, SFX)(P);
    for (k = 0; k < P->nsou; k++)
        for (int a = 0; a < 2; a++)
            CAT(find_i_d_, SFX)(&P->sou[k].icoord[a], &P->sou[k].dcoord[a], P->sou[k].coord[a],
                                a == 0 ? P->oz : P->oxw, a == 0 ? P->dz : P->dx);
    for (k = 0; k < P->ndat; k++)
        for (int a = 0; a < 2; a++)
            CAT(find_i_d_, SFX)(&P->dat[k].icoord[a], &P->dat[k].dcoord[a], P->dat[k].coord[a],
                                a == 0 ? P->oz : P->oxw, a == 0 ? P->dz : P->dx);
    P->wcounter = 0;
    for (it = P->tmin; P->tstep > 0 ? it <= P->tmax : it >= P->tmax; it += P->tstep) {
        if (P->extract >= 0 && P->ndat) CAT(extract_, SFX)(P, buf[3], it);
        if (P->wave) CAT(extract_wave_, SFX)(P, buf[3], it);
        if (P->imaging && P->imaging != 4 && P->tstep == -1) CAT(imaging_, SFX)(P, buf[3], it);
        CAT(free_surface_, SFX)(P, buf[3]);
        if (P->with_rho) CAT(fd_rho_, SFX)(P, buf[3], buf[4]);
        else if (P->stencil == 0) CAT(fd_forward_, SFX)(P, buf[3], buf[4]);
        else CAT(fd_adjoint_, SFX)(P, buf[3], buf[4]);
        if (P->inject && P->nsou) CAT(inject_, SFX)(P, buf[4], it);
        if (P->imaging == 4 && P->tstep == 1) CAT(imaging_, SFX)(P, buf[4], it);
        CAT(time_update_, SFX)(P, buf[0], buf[1], buf[2], buf[3], buf[4]);
        CAT(boundary_, SFX)(P, buf[2], buf[3], buf[4]);
        /* FDswaptimes.f90:14-24 */
        tmp = buf[0]; buf[0] = buf[1]; buf[1] = buf[2]; buf[2] = buf[3]; buf[3] = buf[4]; buf[4] = tmp;
    }
    for (k = 0; k < 5; k++) free(buf[k]);
    free(P->gz); free(P->gx);
    P->gz = P->gx = NULL;
    free(P->s2dzz); free(P->s2dxx);
    P->s2dzz = P->s2dxx = NULL;
}

/* ------------------------------------------------------------------ */
/* Shot window: ExtractPadModel.f90:19-85 extract_coord_source_receiver_patch
 * and the nxw/oxw arithmetic of read_window_vel (:183-185). */
void FN(shot_window)(const sfd_grid *g, const sfd_shot *s, REAL *oxw, int *nxw)
{
    REAL sx = s->sx, gx1 = s->rx[0], gxn = s->rx[0], minx, maxx, mx1, mxn, cx, endx;
    int i;
    for (i = 1; i < s->ntr; i++) {
        gx1 = gx1 < (REAL)s->rx[i] ? gx1 : (REAL)s->rx[i];
        gxn = gxn > (REAL)s->rx[i] ? gxn : (REAL)s->rx[i];
    }
    minx = sx < gx1 ? sx : gx1;
    maxx = sx > gxn ? sx : gxn;
    mx1 = (gx1 + sx) / 2;
    mxn = (gxn + sx) / 2;
    cx = mx1 + FABS(mxn - mx1) / 2;
    *oxw = minx < cx - (REAL)g->aperture_x / 2 ? minx : cx - (REAL)g->aperture_x / 2;
    endx = maxx > cx + (REAL)g->aperture_x / 2 ? maxx : cx + (REAL)g->aperture_x / 2;
    *nxw = NINT((endx - *oxw) / (REAL)g->dx) + 1;
}

/* mod_window_pad(adj=.true.) ExtractPadModel.f90:509-530, then model_pad
 * ReadSouVelRho.f90:545-596 into the padded frame of bounds b. */
void FN(window_pad)(const sfd_grid *g, const REAL *big, REAL oxw, int nxw, const ORC_BOUNDS *b, REAL *vel)
{
    int nz = g->nz, i, j;
    REAL dx = g->dx, ox = g->ox;
    for (j = 1; j <= nxw; j++) {
        REAL cxx = (REAL)(j - 1) * dx + oxw;
        int ix = NINT((cxx - ox) / dx) + 1;
        ix = ix < 1 ? 1 : (ix > g->nx ? g->nx : ix);
        for (i = 1; i <= nz; i++) vel[ORC_IDX(b, i, j)] = (REAL)0 + big[(i - 1) + (size_t)(ix - 1) * nz];
    }
    for (j = 1; j <= nxw; j++) {
        for (i = b->nmin1 - 4; i <= 0; i++) vel[ORC_IDX(b, i, j)] = vel[ORC_IDX(b, 1, j)];
        for (i = nz + 1; i <= b->nmax1 + 4; i++) vel[ORC_IDX(b, i, j)] = vel[ORC_IDX(b, nz, j)];
    }
    for (j = b->nmin2 - 4; j <= 0; j++)
        for (i = b->nmin1 - 4; i <= b->nmax1 + 4; i++) vel[ORC_IDX(b, i, j)] = vel[ORC_IDX(b, i, 1)];
    for (j = nxw + 1; j <= b->nmax2 + 4; j++)
        for (i = b->nmin1 - 4; i <= b->nmax1 + 4; i++) vel[ORC_IDX(b, i, j)] = vel[ORC_IDX(b, i, nxw)];
}

/* mod_copy_image ExtractPadModel.f90:662-695: nearest-index scatter-add of
 * the window image into the full model, skipping out-of-model columns. */
void FN(copy_image)(const sfd_grid *g, REAL oxw, int nxw, const REAL *small, REAL *big)
{
    int i, j;
    for (j = 1; j <= nxw; j++) {
        REAL cxx = (REAL)(j - 1) * (REAL)g->dx + oxw;
        int ix = NINT((cxx - (REAL)g->ox) / (REAL)g->dx) + 1;
        if (ix < 1 || ix > g->nx) continue;
        for (i = 1; i <= g->nz; i++)
            big[(i - 1) + (size_t)(ix - 1) * g->nz] =
                big[(i - 1) + (size_t)(ix - 1) * g->nz] + small[(i - 1) + (size_t)(j - 1) * g->nz];
    }
}

/* Set up one shot's propagation frame: window, bounds, padded model, traces. */
static int CAT(setup_shot_, SFX)(PROP *P, const sfd_grid *g, const REAL *vel_full, const sfd_shot *s)
{
    memset(P, 0, sizeof(*P));
    FN(shot_window)(g, s, &P->oxw, &P->nxw);
    P->nz = g->nz;
    P->oz = g->oz; P->dz = g->dz; P->dx = g->dx; P->dt = g->dt;
    P->lsinc = g->lsinc; P->snapi = g->snapi;
    P->ntsnap = s->nt / g->snapi;
    orc_bounds_t(&P->b, g->nz, P->nxw, g->ntaper, g->lsinc, g->shot_type, g->surf_type);
    P->vel = (REAL *)calloc((size_t)P->b.NZ * P->b.NX, sizeof(REAL));
    if (!P->vel) return -2;
    FN(window_pad)(g, vel_full, P->oxw, P->nxw, &P->b, P->vel);
    P->extract = -1;
    P->rec_type = g->rec_type;
    P->shot_type = g->shot_type;
    P->surf_type = g->surf_type;
    if (g->surf_type != 0) {
        /* ModelSpace_elevation_parameters (ModelSpaceTypes.f90:128-141): find_i_d
         * of each FD column's elevation; columns off the window take the
         * nearest window column's, and that column the nearest model column's */
        const int n2 = P->b.nmax2 - P->b.nmin2 + 1;
        P->ielev = (int *)calloc((size_t)n2, sizeof(int));
        P->delev = (REAL *)calloc((size_t)n2, sizeof(REAL));
        for (int j = P->b.nmin2; j <= P->b.nmax2; j++) {
            int jw = j < 1 ? 1 : (j > P->nxw ? P->nxw : j);
            REAL cxx = (REAL)(jw - 1) * (REAL)g->dx + P->oxw;
            int ix = NINT((cxx - (REAL)g->ox) / (REAL)g->dx) + 1;
            ix = ix < 1 ? 1 : (ix > g->nx ? g->nx : ix);
            REAL e = (orc_elev && ix <= orc_elev_n) ? (REAL)orc_elev[ix - 1] : (REAL)0;
            CAT(find_i_d_, SFX)(&P->ielev[j - P->b.nmin2], &P->delev[j - P->b.nmin2], e, P->oz, P->dz);
            if (P->ielev[j - P->b.nmin2] < 1 || P->ielev[j - P->b.nmin2] > g->nz) return -1;
        }
    }
    if (orc_rho && orc_rho_n == (size_t)g->nz * g->nx) {
        /* read_window_vel's rho: windowed and padded like vel, then Interpolate
         * (ExtractPadModel.f90:259-270) */
        const size_t nb = (size_t)P->b.NZ * P->b.NX, nm = (size_t)g->nz * g->nx;
        REAL *big = (REAL *)malloc(sizeof(REAL) * nm);
        for (size_t q = 0; q < nm; q++) big[q] = (REAL)orc_rho[q];
        P->with_rho = 1;
        P->rho = (REAL *)calloc(nb, sizeof(REAL));
        P->delp = (REAL *)calloc(nb, sizeof(REAL));
        FN(window_pad)(g, big, P->oxw, P->nxw, &P->b, P->rho);
        free(big);
        CAT(rho_half_, SFX)(P);
    }
    return 0;
}

static TR *CAT(make_traces_, SFX)(int n, const float *cz, const float *cx, int nt)
{
    TR *t = (TR *)calloc((size_t)(n > 0 ? n : 1), sizeof(TR));
    for (int k = 0; k < n; k++) {
        t[k].coord[0] = cz[k];
        t[k].coord[1] = cx[k];
        t[k].trace = (REAL *)calloc((size_t)nt, sizeof(REAL));
    }
    return t;
}
static void CAT(free_traces_, SFX)(TR *t, int n)
{
    for (int k = 0; k < n; k++) free(t[k].trace);
    free(t);
}

/* readsou with Born/LSRTM (ReadSouVelRho.f90:318-326): second difference */
static void CAT(second_diff_, SFX)(REAL *tr, int nt, REAL dt)
{
    REAL *o = (REAL *)malloc(sizeof(REAL) * nt);
    memcpy(o, tr, sizeof(REAL) * nt);
    for (int it = 1; it <= nt; it++) tr[it - 1] = 0;
    for (int it = 2; it <= nt - 1; it++)
        tr[it - 1] = tr[it - 1] + ((REAL)-2 * o[it - 1] + o[it - 2] + o[it]) / (dt * dt);
    free(o);
}

/* ---- drivers ------------------------------------------------------ */

/* progs/modeling/Acoustic_modeling_sep.f90:66-88 (and AMOD with extract=NEAREST):
 * forward modeling of one shot.  traces: (nt, ntr) time fastest.
 * snaps (optional, nz*nxw*ntsnap) receives Extraction_wavefield output. */
int FN(modeling)(const sfd_grid *g, const REAL *vel, const sfd_shot *s, const REAL *wavelet,
                 REAL *traces, int extract_kind, REAL *snaps)
{
    PROP P;
    float zz = s->sz, xx = s->sx;
    if (CAT(setup_shot_, SFX)(&P, g, vel, s)) return -2;
    P.stencil = 0;
    P.inject = 1;
    P.nsou = 1;
    P.sou = CAT(make_traces_, SFX)(1, &zz, &xx, s->nt);
    for (int it = 0; it < s->nt; it++) P.sou[0].trace[it] = wavelet[it];
    if (P.with_rho) CAT(half_step_, SFX)(P.sou[0].trace, s->nt, P.lsinc);
    P.extract = extract_kind;
    P.ndat = s->ntr;
    P.dat = CAT(make_traces_, SFX)(s->ntr, s->rz, s->rx, s->nt);
    P.wave = snaps;
    P.tmin = 1; P.tmax = s->nt; P.tstep = 1;
    CAT(propagate_, SFX)(&P);
    for (int r = 0; r < s->ntr; r++)
        memcpy(traces + (size_t)r * s->nt, P.dat[r].trace, sizeof(REAL) * s->nt);
    CAT(free_traces_, SFX)(P.sou, 1);
    CAT(free_traces_, SFX)(P.dat, s->ntr);
    free(P.vel); free(P.ielev); free(P.delev); free(P.rho); free(P.delp);
    return 0;
}

/* Exact discrete adjoint of FN(modeling) (no reference counterpart): the
 * transpose of the linear map wavelet -> traces of the time loop above
 * (sinc injection with gain v^2, forward stencil, 5-level update, Higdon
 * bands, extraction of any kind).  Cotangents L_k of the levels U_k are
 * accumulated backwards in time; each Higdon overwrite u3(i) <- -(sum g u)
 * becomes "a = G(i); G(i) = 0; every input += -g a" in reverse order.
 * dbar: (nt, ntr) time fastest; sbar: nt samples. */
static void CAT(adj_higdon_, SFX)(PROP *P, REAL *L1, REAL *L2, REAL *G, int i, int j, int axis)
{
    const ORC_BOUNDS *b = &P->b;
    const REAL *g;
    int di = 0, dj = 0;
    switch (axis) {
    case -1: g = P->gz + 16 * (j - b->nmin2); di = 1; break;
    case 1: g = P->gz + 16 * (j - b->nmin2) + 8; di = -1; break;
    case -2: g = P->gx + 16 * (i - b->nmin1); dj = 1; break;
    default: g = P->gx + 16 * (i - b->nmin1) + 8; dj = -1; break;
    }
    const REAL a = AT(P, G, i, j);
    AT(P, G, i, j) = 0;
    AT(P, G, i + di, j + dj) = AT(P, G, i + di, j + dj) - g[0] * a;
    AT(P, G, i + 2 * di, j + 2 * dj) = AT(P, G, i + 2 * di, j + 2 * dj) - g[1] * a;
    AT(P, L2, i, j) = AT(P, L2, i, j) - g[2] * a;
    AT(P, L2, i + di, j + dj) = AT(P, L2, i + di, j + dj) - g[3] * a;
    AT(P, L2, i + 2 * di, j + 2 * dj) = AT(P, L2, i + 2 * di, j + 2 * dj) - g[4] * a;
    AT(P, L1, i, j) = AT(P, L1, i, j) - g[5] * a;
    AT(P, L1, i + di, j + dj) = AT(P, L1, i + di, j + dj) - g[6] * a;
    AT(P, L1, i + 2 * di, j + 2 * dj) = AT(P, L1, i + 2 * di, j + 2 * dj) - g[7] * a;
}

/* The reverse sweep shared by FN(modeling_adjoint) and FN(born_adjoint).
 * P: a frame from setup_shot_ (its padded model is freed here).  sbar
 * (optional) receives the transpose of the source injection; with bsnaps,
 * btap and bmbar given, every step k with k % snapi == 0 also accumulates
 * the transpose of Injection_Born (Imaging.f90:779-783, the term
 * 2 taper m S_k / v added to the stencil value) into bmbar (nz, nxw). */
static int CAT(adjoint_core_, SFX)(PROP P, const sfd_shot *s, const REAL *dbar, REAL *sbar, int extract_kind,
                                   const REAL *bsnaps, const REAL *btap, REAL *bmbar)
{
    float zz = s->sz, xx = s->sx;
    const ORC_BOUNDS *b = &P.b;
    const size_t n = (size_t)b->NZ * b->NX;
    const int l = P.lsinc, minx = (int)(-l * 0.5), maxx = (int)(l * 0.5);
    const REAL div11 = (REAL)1. / (REAL)11, dt2 = P.dt * P.dt;
    REAL *buf[5], *G, *W, sz[64], sx[64], wz[64], wx[64];
    int k, i, j, it, r;
    P.nsou = 1;
    P.sou = CAT(make_traces_, SFX)(1, &zz, &xx, s->nt);
    P.ndat = s->ntr;
    P.dat = CAT(make_traces_, SFX)(s->ntr, s->rz, s->rx, s->nt);
    CAT(init_coefs_, SFX)(&P);
    CAT(higdon_tables_, SFX)(&P);
    for (int a = 0; a < 2; a++)
        CAT(find_i_d_, SFX)(&P.sou[0].icoord[a], &P.sou[0].dcoord[a], P.sou[0].coord[a],
                            a == 0 ? P.oz : P.oxw, a == 0 ? P.dz : P.dx);
    for (r = 0; r < P.ndat; r++)
        for (int a = 0; a < 2; a++)
            CAT(find_i_d_, SFX)(&P.dat[r].icoord[a], &P.dat[r].dcoord[a], P.dat[r].coord[a],
                                a == 0 ? P.oz : P.oxw, a == 0 ? P.dz : P.dx);
    for (k = 0; k < 5; k++) buf[k] = (REAL *)calloc(n, sizeof(REAL));
    W = (REAL *)calloc(n, sizeof(REAL));
    /* buf[0] = L_{k+1} (becomes G), buf[1..4] = L_k .. L_{k-3} */
    TR *so = &P.sou[0];
    REAL vs = AT(&P, P.vel, so->icoord[0], so->icoord[1]), v2s = vs * vs;
    FN(mksinc)(sz, l, so->dcoord[0] / P.dz);
    FN(mksinc)(sx, l, so->dcoord[1] / P.dx);
    if (sbar)
        for (it = 0; it < s->nt; it++) sbar[it] = 0;
    for (k = s->nt; k >= 1; k--) {
        G = buf[0];
        /* extraction at step k+1 read U_{k+1}: G += E^T dbar(k+1) */
        if (k + 1 <= s->nt) {
            for (r = 0; r < P.ndat; r++) {
                TR *d = &P.dat[r];
                REAL db = dbar[(size_t)r * s->nt + k], gg;
                if (extract_kind == SFD_EXTRACT_NEAREST) {
                    AT(&P, G, d->icoord[0], d->icoord[1]) = AT(&P, G, d->icoord[0], d->icoord[1]) + db;
                    continue;
                }
                if (extract_kind == SFD_EXTRACT_SINC_V2) {
                    REAL v = AT(&P, P.vel, d->icoord[0], d->icoord[1]);
                    gg = v * v;
                } else if (extract_kind == SFD_EXTRACT_SINC_LSRTM) {
                    gg = (REAL)1 / (P.dz * P.dx);
                } else {
                    gg = 1;
                }
                FN(mksinc)(wz, l, d->dcoord[0] / P.dz);
                FN(mksinc)(wx, l, d->dcoord[1] / P.dx);
                for (j = minx; j <= maxx; j++)
                    for (i = minx; i <= maxx; i++)
                        AT(&P, G, i + d->icoord[0], j + d->icoord[1]) =
                            AT(&P, G, i + d->icoord[0], j + d->icoord[1]) + gg * db * wz[maxx + i] * wx[maxx + j];
            }
        }
        /* Higdon of step k (z-bands, then x-bands), reversed: x-bands first */
        for (j = b->nmax2 + 4; j >= b->nmax2 - 3; j--)
            for (i = b->nmin1; i <= b->nmax1; i++) CAT(adj_higdon_, SFX)(&P, buf[2], buf[1], G, i, j, 2);
        for (j = b->nmin2 - 4; j <= b->nmin2 + 3; j++)
            for (i = b->nmin1; i <= b->nmax1; i++) CAT(adj_higdon_, SFX)(&P, buf[2], buf[1], G, i, j, -2);
        for (j = b->nmin2; j <= b->nmax2; j++) {
            for (i = b->nmax1 + 4; i >= b->nmax1 - 3; i--) CAT(adj_higdon_, SFX)(&P, buf[2], buf[1], G, i, j, 1);
            for (i = b->nmin1 - 4; i <= b->nmin1 + 3; i++) CAT(adj_higdon_, SFX)(&P, buf[2], buf[1], G, i, j, -1);
        }
        /* time update of step k: U_{k+1} = (20U_k - 6U_{k-1} - 4U_{k-2} + U_{k-3} + 12dt^2 (v^2 lap U_k + inj))/11 */
        memset(buf[4], 0, n * sizeof(REAL));   /* L_{k-3}: its first contribution is this step's */
        memset(W, 0, n * sizeof(REAL));
        for (j = b->nmin2; j <= b->nmax2; j++)
            for (i = b->nmin1; i <= b->nmax1; i++) {
                size_t q = ORC_IDX(b, i, j);
                REAL gq = G[q];
                buf[1][q] = buf[1][q] + (REAL)20. * div11 * gq;
                buf[2][q] = buf[2][q] - (REAL)6. * div11 * gq;
                buf[3][q] = buf[3][q] - (REAL)4. * div11 * gq;
                buf[4][q] = buf[4][q] + div11 * gq;
                REAL v = AT(&P, P.vel, i, j);
                W[q] = (REAL)12. * dt2 * div11 * gq * v * v;
            }
        /* the stencil's transpose: L_k(c') += sum over FD cells c of c_tap W(c) */
        for (j = b->nmin2; j <= b->nmax2; j++)
            for (i = b->nmin1; i <= b->nmax1; i++) {
                REAL w = AT(&P, W, i, j);
                if (w == 0) continue;
                AT(&P, buf[1], i, j) = AT(&P, buf[1], i, j) + (P.c[0] + P.c[5]) * w;
                for (int t = 1; t <= 4; t++) {
                    AT(&P, buf[1], i + t, j) = AT(&P, buf[1], i + t, j) + P.c[t] * w;
                    AT(&P, buf[1], i - t, j) = AT(&P, buf[1], i - t, j) + P.c[t] * w;
                    AT(&P, buf[1], i, j + t) = AT(&P, buf[1], i, j + t) + P.c[5 + t] * w;
                    AT(&P, buf[1], i, j - t) = AT(&P, buf[1], i, j - t) + P.c[5 + t] * w;
                }
            }
        /* Injection_Born's transpose: mbar += 12 dt^2/11 G 2 taper S_k / v */
        if (bmbar && k % P.snapi == 0 && k / P.snapi <= P.ntsnap) {
            const size_t nw = (size_t)P.nz * P.nxw;
            const REAL *S = bsnaps + (size_t)(k / P.snapi - 1) * nw;
            for (j = 1; j <= P.nxw; j++)
                for (i = 1; i <= P.nz; i++) {
                    const size_t q = (i - 1) + (size_t)(j - 1) * P.nz;
                    bmbar[q] = bmbar[q] + (REAL)12. * dt2 * div11 * AT(&P, G, i, j) *
                                              ((REAL)2 * btap[j - 1] * S[q] / AT(&P, P.vel, i, j));
                }
        }
        /* the injection's transpose: sbar(k) = 12 dt^2/11 sum v2r wz wx G */
        if (sbar) {
            REAL acc = 0;
            for (j = minx; j <= maxx; j++)
                for (i = minx; i <= maxx; i++)
                    acc = acc + v2s * sz[maxx + i] * sx[maxx + j] *
                                    AT(&P, G, i + so->icoord[0], j + so->icoord[1]);
            sbar[k - 1] = (REAL)12. * dt2 * div11 * acc;
        }
        /* rotate: L_k becomes the next step's L_{k+1} */
        REAL *tmp = buf[0];
        buf[0] = buf[1]; buf[1] = buf[2]; buf[2] = buf[3]; buf[3] = buf[4]; buf[4] = tmp;
    }
    for (k = 0; k < 5; k++) free(buf[k]);
    free(W);
    free(P.gz); free(P.gx);
    CAT(free_traces_, SFX)(P.sou, 1);
    CAT(free_traces_, SFX)(P.dat, s->ntr);
    free(P.vel); free(P.ielev); free(P.delev); free(P.rho); free(P.delp);
    return 0;
}

int FN(modeling_adjoint)(const sfd_grid *g, const REAL *vel, const sfd_shot *s, const REAL *dbar, REAL *sbar,
                         int extract_kind)
{
    if (orc_rho) return -4;   /* variable density: forward modeling, RTM and afwi only */
    if (g->rec_type || g->shot_type || g->surf_type) return -4;   /* no free surface or ghosts */
    PROP P;
    if (CAT(setup_shot_, SFX)(&P, g, vel, s)) return -2;
    return CAT(adjoint_core_, SFX)(P, s, dbar, sbar, extract_kind, NULL, NULL, NULL);
}

/* Exact discrete adjoint of FN(born) (no reference counterpart: the
 * reference's LSRTM imaging is the reverse-time approximation of it).  The
 * map refl -> traces is linear with the background pass fixed: pass 1 is
 * rerun to get the snapshots, the reverse sweep with the LSRTM extraction's
 * transpose accumulates the window's mbar, and the transpose of the window's
 * nearest-column gather (FN(window_pad), edge columns clamped) adds it into
 * mbar (nz, nx).  mbar is accumulated with +=. */
int FN(born_adjoint)(const sfd_grid *g, const REAL *vel, const sfd_shot *s, const REAL *wavelet,
                     const REAL *dbar, REAL *mbar)
{
    if (orc_rho) return -4;   /* variable density: forward modeling, RTM and afwi only */
    if (g->rec_type || g->shot_type || g->surf_type) return -4;   /* no free surface or ghosts */
    PROP P;
    float zz = s->sz, xx = s->sx;
    if (CAT(setup_shot_, SFX)(&P, g, vel, s)) return -2;
    const size_t n = (size_t)P.nz * P.nxw;
    REAL *snaps = (REAL *)calloc(n * (P.ntsnap > 0 ? P.ntsnap : 1), sizeof(REAL));
    REAL *mw = (REAL *)calloc(n, sizeof(REAL));
    REAL *tap = (REAL *)malloc(sizeof(REAL) * P.nxw);
    if (!snaps || !mw || !tap) return -2;
    P.stencil = 0; P.inject = 1; P.nsou = 1;
    P.sou = CAT(make_traces_, SFX)(1, &zz, &xx, s->nt);
    for (int it = 0; it < s->nt; it++) P.sou[0].trace[it] = wavelet[it];
    CAT(second_diff_, SFX)(P.sou[0].trace, s->nt, P.dt);
    P.wave = snaps;
    P.tmin = 1; P.tmax = s->nt; P.tstep = 1;
    CAT(propagate_, SFX)(&P);
    CAT(free_traces_, SFX)(P.sou, 1);
    P.sou = NULL; P.nsou = 0; P.inject = 0; P.wave = NULL;
    FN(compute_taper)(tap, P.nxw);
    const REAL oxw = P.oxw;
    const int nxw = P.nxw;
    int rc = CAT(adjoint_core_, SFX)(P, s, dbar, NULL, SFD_EXTRACT_SINC_LSRTM, snaps, tap, mw);
    for (int j = 1; rc == 0 && j <= nxw; j++) {
        REAL cxx = (REAL)(j - 1) * (REAL)g->dx + oxw;
        int ix = NINT((cxx - (REAL)g->ox) / (REAL)g->dx) + 1;
        ix = ix < 1 ? 1 : (ix > g->nx ? g->nx : ix);
        for (int i = 1; i <= g->nz; i++)
            mbar[(i - 1) + (size_t)(ix - 1) * g->nz] =
                mbar[(i - 1) + (size_t)(ix - 1) * g->nz] + mw[(i - 1) + (size_t)(j - 1) * g->nz];
    }
    free(snaps); free(mw); free(tap);
    return rc;
}

/* progs/rtm/to_disk_to_memory_RTM.f90:358-492 RTM_to_memory (lsrtm=0) and
 * :494-579 LSRTM_to_memory (lsrtm=1), then mod_copy_image into image/illum
 * (full model, accumulated with +=).  vel_bwd may be NULL.  capture
 * (optional) receives the backward u at each imaging step, in imaging order.
 * win_image (optional, nz*nxw) receives the window image. */
int FN(rtm)(const sfd_grid *g, const REAL *vel, const REAL *vel_bwd, const sfd_shot *s,
            const REAL *wavelet, const REAL *traces_in, REAL *image, REAL *illum, int lsrtm,
            REAL *capture, REAL *win_image, REAL *win_illum)
{
    if (orc_rho && lsrtm) return -4;   /* variable density: forward modeling, RTM and afwi only */
    PROP P;
    float zz = s->sz, xx = s->sx;
    if (CAT(setup_shot_, SFX)(&P, g, vel, s)) return -2;
    size_t n = (size_t)P.nz * P.nxw;
    REAL *snaps = (REAL *)calloc(n * (P.ntsnap > 0 ? P.ntsnap : 1), sizeof(REAL));
    REAL *img = (REAL *)calloc(n, sizeof(REAL)), *ill = (REAL *)calloc(n, sizeof(REAL));
    REAL *tap = (REAL *)malloc(sizeof(REAL) * P.nxw);
    if (!snaps || !img || !ill || !tap) return -2;
    /* forward */
    P.stencil = 0; P.inject = 1; P.nsou = 1;
    P.sou = CAT(make_traces_, SFX)(1, &zz, &xx, s->nt);
    for (int it = 0; it < s->nt; it++) P.sou[0].trace[it] = wavelet[it];
    if (lsrtm) CAT(second_diff_, SFX)(P.sou[0].trace, s->nt, P.dt);
    if (P.with_rho) CAT(half_step_, SFX)(P.sou[0].trace, s->nt, P.lsinc);
    P.wave = snaps;
    P.tmin = 1; P.tmax = s->nt; P.tstep = 1;
    CAT(propagate_, SFX)(&P);
    CAT(free_traces_, SFX)(P.sou, 1);
    /* backward */
    if (vel_bwd) FN(window_pad)(g, vel_bwd, P.oxw, P.nxw, &P.b, P.vel);
    FN(compute_taper)(tap, P.nxw);
    P.wave = NULL;
    P.stencil = 1;
    P.inject = lsrtm ? 3 : 1;
    P.nsou = s->ntr;
    P.sou = CAT(make_traces_, SFX)(s->ntr, s->rz, s->rx, s->nt);
    for (int r = 0; r < s->ntr; r++)
        for (int it = 0; it < s->nt; it++) P.sou[r].trace[it] = traces_in[(size_t)r * s->nt + it];
    P.imaging = lsrtm ? 3 : 1;
    P.image = img; P.illum = ill; P.snaps = snaps; P.taperx = tap; P.icounter = 0;
    P.capture = capture; P.ncapture = P.ntsnap;
    P.tmin = s->nt; P.tmax = 1; P.tstep = -1;
    CAT(propagate_, SFX)(&P);
    if (image) FN(copy_image)(g, P.oxw, P.nxw, img, image);
    if (illum) FN(copy_image)(g, P.oxw, P.nxw, ill, illum);
    if (win_image) memcpy(win_image, img, n * sizeof(REAL));
    if (win_illum) memcpy(win_illum, ill, n * sizeof(REAL));
    CAT(free_traces_, SFX)(P.sou, s->ntr);
    free(snaps); free(img); free(ill); free(tap); free(P.vel); free(P.ielev); free(P.delev); free(P.rho); free(P.delp);
    return 0;
}

/* progs/modeling/to_disk_to_memory_BORN.f90:123-220 BORN_to_memory.
 * refl: full-model reflectivity (nz,nx), windowed and padded like vel
 * (readref, ReadSouVelRho.f90:521-543 + mod_window_pad). */
int FN(born)(const sfd_grid *g, const REAL *vel, const REAL *refl, const sfd_shot *s,
             const REAL *wavelet, REAL *traces)
{
    if (orc_rho) return -4;   /* variable density: forward modeling, RTM and afwi only */
    PROP P;
    float zz = s->sz, xx = s->sx;
    if (CAT(setup_shot_, SFX)(&P, g, vel, s)) return -2;
    size_t n = (size_t)P.nz * P.nxw;
    REAL *snaps = (REAL *)calloc(n * (P.ntsnap > 0 ? P.ntsnap : 1), sizeof(REAL));
    REAL *m = (REAL *)calloc((size_t)P.b.NZ * P.b.NX, sizeof(REAL));
    REAL *mw = (REAL *)calloc(n, sizeof(REAL));
    REAL *tap = (REAL *)malloc(sizeof(REAL) * P.nxw);
    P.stencil = 0; P.inject = 1; P.nsou = 1;
    P.sou = CAT(make_traces_, SFX)(1, &zz, &xx, s->nt);
    for (int it = 0; it < s->nt; it++) P.sou[0].trace[it] = wavelet[it];
    CAT(second_diff_, SFX)(P.sou[0].trace, s->nt, P.dt);
    P.wave = snaps;
    P.tmin = 1; P.tmax = s->nt; P.tstep = 1;
    CAT(propagate_, SFX)(&P);
    CAT(free_traces_, SFX)(P.sou, 1);
    FN(window_pad)(g, refl, P.oxw, P.nxw, &P.b, m);
    for (int j = 1; j <= P.nxw; j++)
        for (int i = 1; i <= P.nz; i++) mw[(i - 1) + (size_t)(j - 1) * P.nz] = m[ORC_IDX(&P.b, i, j)];
    FN(compute_taper)(tap, P.nxw);
    P.wave = NULL; P.nsou = 0; P.inject = 0; P.sou = NULL;
    P.extract = SFD_EXTRACT_SINC_LSRTM;
    P.ndat = s->ntr;
    P.dat = CAT(make_traces_, SFX)(s->ntr, s->rz, s->rx, s->nt);
    P.imaging = 4; P.born_m = mw; P.snaps = snaps; P.taperx = tap; P.icounter = 0;
    P.tmin = 1; P.tmax = s->nt; P.tstep = 1;
    CAT(propagate_, SFX)(&P);
    for (int r = 0; r < s->ntr; r++) memcpy(traces + (size_t)r * s->nt, P.dat[r].trace, sizeof(REAL) * s->nt);
    CAT(free_traces_, SFX)(P.dat, s->ntr);
    free(snaps); free(m); free(mw); free(tap); free(P.vel); free(P.ielev); free(P.delev); free(P.rho); free(P.delp);
    return 0;
}

/* Check helper: one interior time step — FD_scheme's forward stencil
 * (FDderivatives.f90:715-743) then the 5-level TimeDer update
 * (FDderivatives.f90:974-976) — on the FD region of a frame with nz x nxw
 * interior cells and ntaper taper cells, with no injection and no absorbing
 * boundary.  um, u0, u1, u2 (u_{-1} .. u2), vel and u3 are padded arrays of
 * the frame (orc_bounds: NZ x NX, z fastest); u3 is written on FD cells. */
int FN(test_step)(const sfd_grid *g, int nxw, const REAL *vel, const REAL *um, const REAL *u0, const REAL *u1,
                  const REAL *u2, REAL *u3)
{
    PROP P;
    memset(&P, 0, sizeof P);
    P.nz = g->nz; P.nxw = nxw;
    P.dz = g->dz; P.dx = g->dx; P.dt = g->dt;
    orc_bounds(&P.b, g->nz, nxw, g->ntaper);
    P.vel = (REAL *)vel;
    CAT(init_coefs_, SFX)(&P);
    CAT(fd_forward_, SFX)(&P, u2, u3);
    CAT(time_update_, SFX)(&P, um, u0, u1, u2, u3);
    return 0;
}

/* Check helper: the AGRAD backward pass (forward stencil, Injection_simple of
 * the traces, reverse time) with u2 captured at every imaging step. */
int FN(test_agrad_capture)(const sfd_grid *g, const REAL *vel, const sfd_shot *s, const REAL *data,
                           REAL *capture)
{
    PROP P;
    if (CAT(setup_shot_, SFX)(&P, g, vel, s)) return -2;
    size_t n = (size_t)P.nz * P.nxw;
    REAL *snaps = (REAL *)calloc(n * (P.ntsnap > 0 ? P.ntsnap : 1), sizeof(REAL));
    REAL *img = (REAL *)calloc(n, sizeof(REAL)), *ill = (REAL *)calloc(n, sizeof(REAL));
    REAL *tap = (REAL *)malloc(sizeof(REAL) * P.nxw);
    FN(compute_taper)(tap, P.nxw);
    P.stencil = 0; P.inject = 2;
    P.nsou = s->ntr;
    P.sou = CAT(make_traces_, SFX)(s->ntr, s->rz, s->rx, s->nt);
    for (int r = 0; r < s->ntr; r++)
        for (int it = 0; it < s->nt; it++) P.sou[r].trace[it] = data[(size_t)r * s->nt + it];
    P.imaging = 1; P.image = img; P.illum = ill; P.snaps = snaps; P.taperx = tap;
    P.capture = capture; P.ncapture = P.ntsnap;
    P.tmin = s->nt; P.tmax = 1; P.tstep = -1;
    CAT(propagate_, SFX)(&P);
    CAT(free_traces_, SFX)(P.sou, s->ntr);
    free(snaps); free(img); free(ill); free(tap); free(P.vel); free(P.ielev); free(P.delev); free(P.rho); free(P.delp);
    return 0;
}

/* Smoothing.f90:192-227 boxconv */
static void CAT(boxconv_, SFX)(int nb, int nx, const REAL *xx, REAL *yy)
{
    int ny = nx + nb - 1, i;
    REAL *bb = (REAL *)calloc((size_t)(nx + nb), sizeof(REAL));
    bb[0] = xx[0];
    for (i = 2; i <= nx; i++) bb[i - 1] = bb[i - 2] + xx[i - 1];
    for (i = nx + 1; i <= ny; i++) bb[i - 1] = bb[i - 2];
    for (i = 1; i <= nb; i++) yy[i - 1] = bb[i - 1];
    for (i = nb + 1; i <= ny; i++) yy[i - 1] = bb[i - 1] - bb[i - nb - 1];
    for (i = 1; i <= ny; i++) yy[i - 1] = yy[i - 1] / (REAL)nb;
    free(bb);
}

/* Smoothing.f90:140-181 triangle: in/out strided vectors */
void FN(triangle)(int nr, int n12, const REAL *uu, size_t su, REAL *vv, size_t sv)
{
    REAL *pp = (REAL *)calloc((size_t)(n12 + nr), sizeof(REAL));
    REAL *qq = (REAL *)calloc((size_t)(n12 + 2 * nr), sizeof(REAL));
    REAL *tt = (REAL *)calloc((size_t)n12, sizeof(REAL));
    int i, np;
    for (i = 1; i <= n12; i++) qq[i - 1] = uu[(i - 1) * su];
    if (n12 == 1) {
        tt[0] = qq[0];
    } else {
        CAT(boxconv_, SFX)(nr, n12, qq, pp);
        np = nr + n12 - 1;
        CAT(boxconv_, SFX)(nr, np, pp, qq);
        for (i = 1; i <= n12; i++) tt[i - 1] = qq[i + nr - 2];
        for (i = 1; i <= nr - 1; i++) tt[i - 1] = tt[i - 1] + qq[nr - i - 1];
        for (i = 1; i <= nr - 1; i++) tt[n12 - i] = tt[n12 - i] + qq[n12 + (nr - 1) + i - 1];
    }
    for (i = 1; i <= n12; i++) vv[(i - 1) * sv] = tt[i - 1];
    free(pp); free(qq); free(tt);
}

/* Smoothing.f90:87-138 triangle2, 2D (n3 = 1) */
void FN(triangle2)(int n1, int n2, int rect1, int rect2, int doz, int dox, REAL *uu)
{
    size_t n = (size_t)n1 * n2;
    REAL *ss = (REAL *)calloc(n, sizeof(REAL));
    int i1, i2;
    if (dox) {
        for (i1 = 0; i1 < n1; i1++) FN(triangle)(rect2, n2, uu + i1, (size_t)n1, ss + i1, (size_t)n1);
        memcpy(uu, ss, n * sizeof(REAL));
    }
    if (doz) {
        for (i2 = 0; i2 < n2; i2++) FN(triangle)(rect1, n1, uu + (size_t)i2 * n1, 1, ss + (size_t)i2 * n1, 1);
        memcpy(uu, ss, n * sizeof(REAL));
    }
    free(ss);
}

/* Norm.f90:12-79 fct_compute on one trace rd(nd).  *thresh is the caller's
 * THRESH, overwritten with max|rd| when it is 0 (Huber, Cauchy, hyperbolic;
 * :36-40,52-56,65-69).  The L2 sum is of FP64 products (dprod); the others sum
 * their terms in REAL, in order, as SUM of a default-real array does. */
double FN(fct_compute)(int ntype, const REAL *rd, int nd, REAL *thresh)
{
    if (ntype == 2 || ntype == 0) {
        double f = 0;
        for (int i = 0; i < nd; i++) f += (double)rd[i] * (double)rd[i];
        return f / 2;
    }
    if (ntype != 1 && *thresh == (REAL)0) {
        REAL m = 0;
        for (int i = 0; i < nd; i++) m = FABS(rd[i]) > m ? FABS(rd[i]) : m;
        *thresh = m;
    }
    const REAL t = *thresh;
    REAL acc = 0;
    for (int i = 0; i < nd; i++) {
        const REAL r = rd[i], a = FABS(r);
        REAL term;
        if (ntype == 1) term = a;
        else if (ntype == 12) term = a > t ? a - t / (REAL)2 : r * r / ((REAL)2 * t);
        else if (ntype == 3) { const REAL q = r / t; term = LOG((REAL)1 + q * q) / (REAL)2; }
        else term = SQRT(t * t + r * r) - t;
        acc = acc + term;
    }
    return (double)acc;
}

/* Norm.f90:81-141 gdt_compute: rd(nd) becomes the adjoint source. */
void FN(gdt_compute)(int ntype, REAL *rd, int nd, REAL t)
{
    for (int i = 0; i < nd; i++) {
        const REAL r = rd[i];
        if (ntype == 1) rd[i] = r < 0 ? (REAL)-1 : (r > 0 ? (REAL)1 : (REAL)0);
        else if (ntype == 12) {   /* max(-1, min(1, q)); a NaN q stays NaN, as the reference's MIN/MAX */
            const REAL q = r / t;
            rd[i] = q > (REAL)1 ? (REAL)1 : (q < (REAL)-1 ? (REAL)-1 : q);
        }
        else if (ntype == 3) rd[i] = r / (t * t + r * r);
        else if (ntype == 4) rd[i] = (r / t) / SQRT((REAL)1 + r * r / (t * t));
    }
}

/* compute_function_gradient.f90:208-350 compute_fct_gdt (single thread, shots
 * summed in order), with Compute_OF_RES_ADJ (Objective_...:17-40) and its
 * norms (Norm.f90:12-141: o->norm, o->norm_thresh) and AMOD/AGRAD (to_disk_to_memory_AMOD.f90:27-82,
 * to_disk_to_memory_AGRAD.f90:27-78).  dobs/resid: per shot (nt,ntr) blocks
 * concatenated; mute same layout (NULL = 1); vpmask (nz*nx, NULL = 1).
 * Shots may differ in nt (ragged batch; the misfit scaling sums each
 * shot's 2 nt ntr).  grad/illu_out are (nz,nx). */
/* progs/afwi/3DAFWI.f90 one-shot gradient, AFWI_to_memory
 * (to_disk_to_memory_AFWI.f90:229-415), 2D, per shot: forward modeling with
 * Injection_sinc (readsou's half-step shift with a density) and
 * Extraction_array_sinc_afwi (gain 1) storing snapshots; Compute_OF_RES_3D
 * (Objective_Function_Residual_Adjsource.f90:42-67: one THRESH, in trace
 * order; the misfit starts at 0 where the driver leaves it unset); reverse
 * time with the adjoint-style stencil (the density operator with a density),
 * Injection_sinc of the residual and Imaging_condition_AFWI_from_memory, or
 * Imaging_condition_AFWI_RHOVP_from_memory with a density; then
 * mod_copy_gradient_to_disk (ExtractPadModel.f90:592-660): each window added
 * into the model, gradients summed over the shots, and with a density and
 * vprho 1 the Vp/Ip change of variables on the model's vel and rho.  grad:
 * (nz,nx) for vp, then (nz,nx) for rho (or Ip) with a density. */
int FN(afwi)(const sfd_grid *g, const REAL *vel, int nshots, const sfd_shot *shots, const REAL *wavelet,
             const REAL *dobs, const REAL *mute, int norm, REAL norm_thresh, int vprho, double *fout, REAL *grad,
             REAL *illum)
{
    const size_t N = (size_t)g->nz * g->nx;
    const int np = orc_rho && orc_rho_n == N ? 2 : 1;
    size_t off = 0;
    double f = 0;
    for (size_t q = 0; q < N * np; q++) grad[q] = 0;
    for (size_t q = 0; q < N; q++) illum[q] = 0;
    for (int is = 0; is < nshots; is++) {
        const sfd_shot *s = &shots[is];
        PROP P;
        float zz = s->sz, xx = s->sx;
        if (CAT(setup_shot_, SFX)(&P, g, vel, s)) return -2;
        if (P.with_rho && P.nxw < 2) return -1;   /* the one-sided x gradient needs two window columns */
        size_t n = (size_t)P.nz * P.nxw;
        REAL *snaps = (REAL *)calloc(n * (P.ntsnap > 0 ? P.ntsnap : 1), sizeof(REAL));
        REAL *img = (REAL *)calloc(n, sizeof(REAL)), *img2 = (REAL *)calloc(n, sizeof(REAL));
        REAL *ill = (REAL *)calloc(n, sizeof(REAL));
        REAL *tap = (REAL *)malloc(sizeof(REAL) * P.nxw);
        if (!snaps || !img || !img2 || !ill || !tap) return -2;
        P.stencil = 0; P.inject = 1; P.nsou = 1;
        P.sou = CAT(make_traces_, SFX)(1, &zz, &xx, s->nt);
        for (int it = 0; it < s->nt; it++) P.sou[0].trace[it] = wavelet[it];
        if (P.with_rho) CAT(half_step_, SFX)(P.sou[0].trace, s->nt, P.lsinc);
        P.extract = SFD_EXTRACT_SINC_UNIT;
        P.ndat = s->ntr;
        P.dat = CAT(make_traces_, SFX)(s->ntr, s->rz, s->rx, s->nt);
        P.wave = snaps;
        P.tmin = 1; P.tmax = s->nt; P.tstep = 1;
        CAT(propagate_, SFX)(&P);
        CAT(free_traces_, SFX)(P.sou, 1);
        REAL thresh = norm_thresh;
        for (int r = 0; r < s->ntr; r++) {
            REAL *d = P.dat[r].trace;
            const REAL *ob = dobs + off + (size_t)r * s->nt;
            const REAL *mu = mute ? mute + off + (size_t)r * s->nt : NULL;
            for (int it = 0; it < s->nt; it++) d[it] = (mu ? mu[it] : (REAL)1) * (ob[it] - d[it]);
            f = f + FN(fct_compute)(norm, d, s->nt, &thresh);
            for (int it = 0; it < s->nt; it++) d[it] = (mu ? mu[it] : (REAL)1) * d[it];
            FN(gdt_compute)(norm, d, s->nt, thresh);
        }
        FN(compute_taper)(tap, P.nxw);
        P.wave = NULL; P.extract = -1;
        P.stencil = 1; P.inject = 1;
        P.nsou = s->ntr; P.sou = P.dat; P.ndat = 0; P.dat = NULL;
        P.imaging = P.with_rho ? 5 : 2;
        P.image = img; P.image2 = img2; P.illum = ill; P.snaps = snaps; P.taperx = tap;
        P.dt2snap = (P.dt * (REAL)P.snapi) * (P.dt * (REAL)P.snapi);
        P.icounter = 0;
        P.tmin = s->nt; P.tmax = 1; P.tstep = -1;
        CAT(propagate_, SFX)(&P);
        FN(copy_image)(g, P.oxw, P.nxw, img, grad);
        if (np == 2) FN(copy_image)(g, P.oxw, P.nxw, img2, grad + N);
        FN(copy_image)(g, P.oxw, P.nxw, ill, illum);
        CAT(free_traces_, SFX)(P.sou, s->ntr);
        free(snaps); free(img); free(img2); free(ill); free(tap);
        free(P.vel); free(P.ielev); free(P.delev); free(P.rho); free(P.delp);
        off += (size_t)s->nt * s->ntr;
    }
    if (np == 2 && vprho == 1)   /* ExtractPadModel.f90:643-646 on the model's vel and rho */
        for (size_t q = 0; q < N; q++) {
            const REAL v = vel[q], r = (REAL)orc_rho[q];
            grad[q] = grad[q] - (r / v) * grad[N + q];
            grad[N + q] = ((REAL)1 / v) * grad[N + q];
        }
    *fout = f;
    return 0;
}

int FN(fwi_gradient)(const sfd_grid *g, const REAL *vel, int nshots, const sfd_shot *shots,
                     const REAL *wavelet, const REAL *dobs, const REAL *mute, const REAL *vpmask,
                     const sfd_fwi_opts *o, double *fout, REAL *grad, REAL *illu_out, REAL *resid)
{
    if (orc_rho) return -4;   /* variable density: forward modeling, RTM and afwi only */
    size_t N = (size_t)g->nz * g->nx, off = 0;
    REAL *illu = (REAL *)calloc(N, sizeof(REAL));
    double f = 0;
    int ntot = 0, n1 = shots[0].nt;
    for (size_t q = 0; q < N; q++) grad[q] = 0;
    for (int is = 0; is < nshots; is++) {
        const sfd_shot *s = &shots[is];
        PROP P;
        float zz = s->sz, xx = s->sx;
        if (CAT(setup_shot_, SFX)(&P, g, vel, s)) return -2;
        size_t n = (size_t)P.nz * P.nxw;
        REAL *snaps = (REAL *)calloc(n * (P.ntsnap > 0 ? P.ntsnap : 1), sizeof(REAL));
        REAL *img = (REAL *)calloc(n, sizeof(REAL)), *ill = (REAL *)calloc(n, sizeof(REAL));
        REAL *tap = (REAL *)malloc(sizeof(REAL) * P.nxw);
        /* AMOD */
        P.stencil = 0; P.inject = 1; P.nsou = 1;
        P.sou = CAT(make_traces_, SFX)(1, &zz, &xx, s->nt);
        for (int it = 0; it < s->nt; it++) P.sou[0].trace[it] = wavelet[it];
        P.extract = SFD_EXTRACT_NEAREST;
        P.ndat = s->ntr;
        P.dat = CAT(make_traces_, SFX)(s->ntr, s->rz, s->rx, s->nt);
        P.wave = snaps;
        P.tmin = 1; P.tmax = s->nt; P.tstep = 1;
        CAT(propagate_, SFX)(&P);
        CAT(free_traces_, SFX)(P.sou, 1);
        /* Compute_OF_RES_ADJ: one THRESH for the shot's traces, in order */
        REAL thresh = (REAL)o->norm_thresh;
        for (int r = 0; r < s->ntr; r++) {
            REAL *d = P.dat[r].trace;
            const REAL *ob = dobs + off + (size_t)r * s->nt;
            const REAL *mu = mute ? mute + off + (size_t)r * s->nt : NULL;
            for (int it = 0; it < s->nt; it++) {
                REAL w = mu ? mu[it] : (REAL)1;
                d[it] = w * (ob[it] - d[it]);
            }
            f = f + FN(fct_compute)(o->norm, d, s->nt, &thresh);
            if (resid) memcpy(resid + off + (size_t)r * s->nt, d, sizeof(REAL) * s->nt);
            for (int it = 0; it < s->nt; it++) d[it] = (mu ? mu[it] : (REAL)1) * d[it];
            FN(gdt_compute)(o->norm, d, s->nt, thresh);
        }
        /* AGRAD */
        FN(compute_taper)(tap, P.nxw);
        P.wave = NULL; P.extract = -1;
        P.stencil = 0; P.inject = 2;
        P.nsou = s->ntr; P.sou = P.dat; P.ndat = 0; P.dat = NULL;
        P.imaging = 2; P.image = img; P.illum = ill; P.snaps = snaps; P.taperx = tap;
        P.dt2snap = (P.dt * (REAL)P.snapi) * (P.dt * (REAL)P.snapi);
        P.icounter = 0;
        P.tmin = s->nt; P.tmax = 1; P.tstep = -1;
        CAT(propagate_, SFX)(&P);
        FN(copy_image)(g, P.oxw, P.nxw, img, grad);
        FN(copy_image)(g, P.oxw, P.nxw, ill, illu);
        CAT(free_traces_, SFX)(P.sou, s->ntr);
        free(snaps); free(img); free(ill); free(tap); free(P.vel); free(P.ielev); free(P.delev); free(P.rho); free(P.delp);
        off += (size_t)s->nt * s->ntr;
        ntot += s->ntr;
    }
    /* :332-343 */
    if (o->illupow != 1.0f)
        for (size_t q = 0; q < N; q++) illu[q] = POW(illu[q], (REAL)o->illupow);
    /* f / (2 n1 ntr) of :332-336 with the reference's one nt; a ragged batch
     * (shots of their own nt) sums 2 nt ntr over its shots, which is the same
     * number when every nt is n1 */
    double scaling = 0;
    for (int is = 0; is < nshots; is++) scaling += 2.0 * (double)shots[is].nt * (double)shots[is].ntr;
    (void)n1; (void)ntot;
    f = f / scaling;
    REAL mx = illu[0];
    double ss = 0;
    for (size_t q = 0; q < N; q++) {
        grad[q] = (REAL)2 * grad[q] / (REAL)scaling;
        mx = illu[q] > mx ? illu[q] : mx;
        ss += (double)illu[q] * (double)illu[q];
    }
    double rn = sqrt(ss / (double)N);
    for (size_t q = 0; q < N; q++) illu[q] = (REAL)((double)(illu[q] + mx / (REAL)10000) / rn);
    for (size_t q = 0; q < N; q++) {
        grad[q] = grad[q] / illu[q];
        grad[q] = grad[q] * (vpmask ? vpmask[q] : (REAL)1);
    }
    FN(triangle2)(g->nz, g->nx, o->rect_z, o->rect_x, o->do_z, o->do_x, grad);
    for (size_t q = 0; q < N; q++) grad[q] = grad[q] * (vpmask ? vpmask[q] : (REAL)1);
    if (illu_out) memcpy(illu_out, illu, N * sizeof(REAL));
    free(illu);
    *fout = f;
    return 0;
}

#undef AT
#undef TR
#undef PROP
#undef NINT
#undef FN
#undef CAT
#undef CAT_
