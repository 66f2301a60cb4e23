/*
 * lbfgs_oracle.c — CPU restatement of the reference's in-process L-BFGS:
 * libs/libinv/lbfgs.f  FREEZE_X (:1-47), LBFGS (:49-558), MCSRCH (:652-982),
 * MCSTEP (:985-1210), block data LB2 (:636-641: GTOL=0.9, STPMIN=1e-20,
 * STPMAX=1e20) and the in-repo DDOT (libs/libinv/blas.f:58-105, unrolled by 5).
 *
 * Also the out-of-core variant libs/libnlinv/nlinv_lbfgs.f: FREEZE_XOC (:1-22),
 * LBFGSOC (:31-599), MCSRCHOC (:614-967), block data LB4 (:981-986, the same
 * GTOL/STPMIN/STPMAX).  It is LBFGS/MCSRCH with a per-parameter clip, its own
 * STP1 rule and the SAVE state written to files between calls; orc_lbfgsoc
 * keeps that state in memory (the engine's file round trip is checked by
 * equality with it).
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the engine's sfd_lbfgs_* and
 * sfd_lbfgsoc_step).
 *
 * The Fortran SAVE variables and GOTO re-entry points become the fields of
 * orc_lbfgs_t and an explicit `entry` switch; the reverse-communication
 * protocol (IFLAG 0 -> 1 -> 1 ..., INFO -1 inside MCSRCH) is unchanged.
 * Arrays are 0-based here: W(k) of the reference is w[k-1].
 * The reference's OpenMP REDUCTION dot products (lbfgs.f:361-463) are summed
 * serially, which is what the reference computes with one thread.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    /* LBFGS SAVE state */
    int iter, nfun, point, ispt, iypt, maxfev, info, bound, npt, cp, nfev, finish;
    double gnorm, xnorm2, stp1, ftol, stp, ys, yy;
    double gtol, stpmin, stpmax;
    /* MCSRCH SAVE state */
    int infoc, brackt, stage1;
    double dginit, dgtest, finit, width, width1, stx, fx, dgx, sty, fy, dgy, stmin, stmax;
} orc_lbfgs_t;

double orc_ddot(int n, const double *dx, const double *dy)
{
    double dtemp = 0.0;
    int i, m;
    if (n <= 0) return 0.0;
    m = n % 5;
    for (i = 0; i < m; i++) dtemp = dtemp + dx[i] * dy[i];
    if (n < 5) return dtemp;
    for (i = m; i < n; i += 5)
        dtemp = dtemp + dx[i] * dy[i] + dx[i + 1] * dy[i + 1] + dx[i + 2] * dy[i + 2] +
                dx[i + 3] * dy[i + 3] + dx[i + 4] * dy[i + 4];
    return dtemp;
}

/* lbfgs.f:1-47 */
void orc_freeze_x(double xmin, double xmax, const float *mask, int type, double *x, int n,
                  const double *xinit)
{
    int i;
    if (type) {
        for (i = 0; i < n; i++) {
            x[i] = fmax(xmin, x[i]);
            x[i] = fmin(xmax, x[i]);
        }
    } else {
        for (i = 0; i < n; i++)
            if (mask[i] != 0 && mask[i] != 2) {
                x[i] = fmax(xmin, x[i]);
                x[i] = fmin(xmax, x[i]);
            }
        for (i = 0; i < n; i++) {
            if (mask[i] == 0) x[i] = xinit[i];
            if (mask[i] == 2) x[i] = xinit[i];
        }
    }
}

/* nlinv_lbfgs.f:1-22 FREEZE_XOC: parameter j's block of M = N/NPARAM values
 * (integer division; a remainder is left alone) clipped to [XMIN(j), XMAX(j)] */
void orc_freeze_xoc(const double *xmin, const double *xmax, int nparam, double *x, int n)
{
    int i, j, m = n / nparam;
    for (j = 0; j < nparam; j++)
        for (i = j * m; i < (j + 1) * m; i++) {
            x[i] = fmax(xmin[j], x[i]);
            x[i] = fmin(xmax[j], x[i]);
        }
}

/* The bound handling of one variant: FREEZE_X (libinv) or FREEZE_XOC (libnlinv). */
typedef struct {
    int oc;
    double xmin, xmax;
    const float *mask;
    int type;
    const double *xinit;
    const double *pxmin, *pxmax;
    int nparam;
    int boundt;
} orc_clip_t;

static void clip(const orc_clip_t *c, double *x, int n)
{
    if (c->oc) orc_freeze_xoc(c->pxmin, c->pxmax, c->nparam, x, n);
    else orc_freeze_x(c->xmin, c->xmax, c->mask, c->type, x, n, c->xinit);
}

/* lbfgs.f:985-1210 MCSTEP */
static void mcstep(double *stx, double *fx, double *dx, double *sty, double *fy, double *dy,
                   double *stp, double fp, double dp, int *brackt, double stpmin, double stpmax,
                   int *info)
{
    int bound;
    double gamma, p, q, r, s, sgnd, stpc, stpf, stpq, theta;
    *info = 0;
    if ((*brackt && (*stp <= fmin(*stx, *sty) || *stp >= fmax(*stx, *sty))) ||
        *dx * (*stp - *stx) >= 0.0 || stpmax < stpmin)
        return;
    sgnd = dp * (*dx / fabs(*dx));
    if (fp > *fx) {
        *info = 1;
        bound = 1;
        theta = 3 * (*fx - fp) / (*stp - *stx) + *dx + dp;
        s = fmax(fabs(theta), fmax(fabs(*dx), fabs(dp)));
        gamma = s * sqrt((theta / s) * (theta / s) - (*dx / s) * (dp / s));
        if (*stp < *stx) gamma = -gamma;
        p = (gamma - *dx) + theta;
        q = ((gamma - *dx) + gamma) + dp;
        r = p / q;
        stpc = *stx + r * (*stp - *stx);
        stpq = *stx + ((*dx / ((*fx - fp) / (*stp - *stx) + *dx)) / 2) * (*stp - *stx);
        if (fabs(stpc - *stx) < fabs(stpq - *stx))
            stpf = stpc;
        else
            stpf = stpc + (stpq - stpc) / 2;
        *brackt = 1;
    } else if (sgnd < 0.0) {
        *info = 2;
        bound = 0;
        theta = 3 * (*fx - fp) / (*stp - *stx) + *dx + dp;
        s = fmax(fabs(theta), fmax(fabs(*dx), fabs(dp)));
        gamma = s * sqrt((theta / s) * (theta / s) - (*dx / s) * (dp / s));
        if (*stp > *stx) gamma = -gamma;
        p = (gamma - dp) + theta;
        q = ((gamma - dp) + gamma) + *dx;
        r = p / q;
        stpc = *stp + r * (*stx - *stp);
        stpq = *stp + (dp / (dp - *dx)) * (*stx - *stp);
        if (fabs(stpc - *stp) > fabs(stpq - *stp))
            stpf = stpc;
        else
            stpf = stpq;
        *brackt = 1;
    } else if (fabs(dp) < fabs(*dx)) {
        *info = 3;
        bound = 1;
        theta = 3 * (*fx - fp) / (*stp - *stx) + *dx + dp;
        s = fmax(fabs(theta), fmax(fabs(*dx), fabs(dp)));
        gamma = s * sqrt(fmax(0.0, (theta / s) * (theta / s) - (*dx / s) * (dp / s)));
        if (*stp > *stx) gamma = -gamma;
        p = (gamma - dp) + theta;
        q = (gamma + (*dx - dp)) + gamma;
        r = p / q;
        if (r < 0.0 && gamma != 0.0)
            stpc = *stp + r * (*stx - *stp);
        else if (*stp > *stx)
            stpc = stpmax;
        else
            stpc = stpmin;
        stpq = *stp + (dp / (dp - *dx)) * (*stx - *stp);
        if (*brackt) {
            if (fabs(*stp - stpc) < fabs(*stp - stpq))
                stpf = stpc;
            else
                stpf = stpq;
        } else {
            if (fabs(*stp - stpc) > fabs(*stp - stpq))
                stpf = stpc;
            else
                stpf = stpq;
        }
    } else {
        *info = 4;
        bound = 0;
        if (*brackt) {
            theta = 3 * (fp - *fy) / (*sty - *stp) + *dy + dp;
            s = fmax(fabs(theta), fmax(fabs(*dy), fabs(dp)));
            gamma = s * sqrt((theta / s) * (theta / s) - (*dy / s) * (dp / s));
            if (*stp > *sty) gamma = -gamma;
            p = (gamma - dp) + theta;
            q = ((gamma - dp) + gamma) + *dy;
            r = p / q;
            stpc = *stp + r * (*sty - *stp);
            stpf = stpc;
        } else if (*stp > *stx) {
            stpf = stpmax;
        } else {
            stpf = stpmin;
        }
    }
    if (fp > *fx) {
        *sty = *stp;
        *fy = fp;
        *dy = dp;
    } else {
        if (sgnd < 0.0) {
            *sty = *stx;
            *fy = *fx;
            *dy = *dx;
        }
        *stx = *stp;
        *fx = fp;
        *dx = dp;
    }
    stpf = fmin(stpmax, stpf);
    stpf = fmax(stpmin, stpf);
    *stp = stpf;
    if (*brackt && bound) {
        if (*sty > *stx)
            /* 0.66 is a default-REAL literal there: its FP32 value, widened (lbfgs.f:1203) */
            *stp = fmin(*stx + (double)0.66f * (*sty - *stx), *stp);
        else
            *stp = fmax(*stx + (double)0.66f * (*sty - *stx), *stp);
    }
}

/* lbfgs.f:652-982 MCSRCH; wa is the reference's DIAG array (passed as WA) */
static void mcsrch(orc_lbfgs_t *S, int n, double *x, double f, const double *g, const double *s,
                   double *stp, double ftol, double xtol, int maxfev, int *info, int *nfev, double *wa,
                   const orc_clip_t *c)
{
    /* XTRAPF: 4.0D1 in libinv's MCSRCH (lbfgs.f:792), 4.0D0 in MCSRCHOC (nlinv_lbfgs.f:752) */
    const double p5 = 0.5, p66 = 0.66, xtrapf = c->oc ? 4.0 : 40.0, zero = 0.0;
    int j;
    double dg, dgm, dgxm, dgym, fm, fxm, fym, ftest1;
    if (*info == -1) goto L45;
    S->infoc = 1;
    if (n <= 0 || *stp <= zero || ftol < zero || S->gtol < zero || xtol < zero || S->stpmin < zero ||
        S->stpmax < S->stpmin || maxfev <= 0)
        return;
    S->dginit = zero;
    for (j = 0; j < n; j++) S->dginit = S->dginit + g[j] * s[j];
    if (S->dginit >= zero) return;
    S->brackt = 0;
    S->stage1 = 1;
    *nfev = 0;
    S->finit = f;
    S->dgtest = ftol * S->dginit;
    S->width = S->stpmax - S->stpmin;
    S->width1 = S->width / p5;
    for (j = 0; j < n; j++) wa[j] = x[j];
    S->stx = zero;
    S->fx = S->finit;
    S->dgx = S->dginit;
    S->sty = zero;
    S->fy = S->finit;
    S->dgy = S->dginit;
L30:
    if (S->brackt) {
        S->stmin = fmin(S->stx, S->sty);
        S->stmax = fmax(S->stx, S->sty);
    } else {
        S->stmin = S->stx;
        S->stmax = *stp + xtrapf * (*stp - S->stx);
    }
    *stp = fmax(*stp, S->stpmin);
    *stp = fmin(*stp, S->stpmax);
    if ((S->brackt && (*stp <= S->stmin || *stp >= S->stmax)) || *nfev >= maxfev - 1 || S->infoc == 0 ||
        (S->brackt && S->stmax - S->stmin <= xtol * S->stmax))
        *stp = S->stx;
    for (j = 0; j < n; j++) x[j] = wa[j] + *stp * s[j];
    if (c->boundt == 1) clip(c, x, n);
    *info = -1;
    return;
L45:
    *info = 0;
    *nfev = *nfev + 1;
    dg = zero;
    for (j = 0; j < n; j++) dg = dg + g[j] * s[j];
    ftest1 = S->finit + *stp * S->dgtest;
    if ((S->brackt && (*stp <= S->stmin || *stp >= S->stmax)) || S->infoc == 0) *info = 6;
    if (*stp == S->stpmax && f <= ftest1 && dg <= S->dgtest) *info = 5;
    if (*stp == S->stpmin && (f > ftest1 || dg >= S->dgtest)) *info = 4;
    if (*nfev >= maxfev) *info = 3;
    if (S->brackt && S->stmax - S->stmin <= xtol * S->stmax) *info = 2;
    if (f <= ftest1 && fabs(dg) <= S->gtol * (-S->dginit)) *info = 1;
    if (*info != 0) return;
    if (S->stage1 && f <= ftest1 && dg >= fmin(ftol, S->gtol) * S->dginit) S->stage1 = 0;
    if (S->stage1 && f <= S->fx && f > ftest1) {
        fm = f - *stp * S->dgtest;
        fxm = S->fx - S->stx * S->dgtest;
        fym = S->fy - S->sty * S->dgtest;
        dgm = dg - S->dgtest;
        dgxm = S->dgx - S->dgtest;
        dgym = S->dgy - S->dgtest;
        mcstep(&S->stx, &fxm, &dgxm, &S->sty, &fym, &dgym, stp, fm, dgm, &S->brackt, S->stmin, S->stmax,
               &S->infoc);
        S->fx = fxm + S->stx * S->dgtest;
        S->fy = fym + S->sty * S->dgtest;
        S->dgx = dgxm + S->dgtest;
        S->dgy = dgym + S->dgtest;
    } else {
        mcstep(&S->stx, &S->fx, &S->dgx, &S->sty, &S->fy, &S->dgy, stp, f, dg, &S->brackt, S->stmin,
               S->stmax, &S->infoc);
    }
    if (S->brackt) {
        if (fabs(S->sty - S->stx) >= p66 * S->width1) *stp = S->stx + p5 * (S->sty - S->stx);
        S->width1 = S->width;
        S->width = fabs(S->sty - S->stx);
    }
    goto L30;
}

orc_lbfgs_t *orc_lbfgs_new(void)
{
    orc_lbfgs_t *S = (orc_lbfgs_t *)calloc(1, sizeof(orc_lbfgs_t));
    S->gtol = 0.9;
    S->stpmin = 1.0e-20;
    S->stpmax = 1.0e20;
    return S;
}
void orc_lbfgs_free(orc_lbfgs_t *S) { free(S); }

/* lbfgs.f:49-558 LBFGS (DIAGCO=.false., IPRINT ignored: output only), and
 * LBFGSOC (nlinv_lbfgs.f:31-599) when c->oc: the two differ only in the clip
 * and in STP1 (below). */
static void lbfgs_core(orc_lbfgs_t *S, int n, int m, double *x, double f, const double *g, double *diag,
                       double eps, double xtol, double *w, int *iflag, int *myinfo, const orc_clip_t *c,
                       int stp1_opt, double step_init)
{
    int i, j, inmc, iycn, iscn;
    double sq, yr, beta, xnorm;
    if (*iflag == 0) goto L10;
    if (*iflag == 1) goto L172;
    if (*iflag == 2) goto L100;
L10:
    S->iter = 0;
    if (n <= 0 || m <= 0) { *iflag = -3; return; }
    if (S->gtol <= 1.0e-04) S->gtol = 9.0e-01;
    S->nfun = 1;
    S->point = 0;
    S->finish = 0;
    for (i = 0; i < n; i++) diag[i] = 1.0;
    S->ispt = n + 2 * m;
    S->iypt = S->ispt + n * m;
    for (i = 1; i <= n; i++) w[S->ispt + i - 1] = -g[i - 1] * diag[i - 1];
    S->gnorm = sqrt(orc_ddot(n, g, g));
    S->xnorm2 = sqrt(orc_ddot(n, x, x));
    if (!c->oc) {                        /* lbfgs.f:333-338 */
        if (c->boundt == 0)
            S->stp1 = 1.0 / S->gnorm;
        else
            S->stp1 = S->xnorm2 / S->gnorm / 100;
    } else {                             /* nlinv_lbfgs.f:336-344 */
        if (stp1_opt == 1)
            S->stp1 = S->xnorm2 / S->gnorm / 100;
        else
            S->stp1 = 1.0 / S->gnorm;
        if (step_init != 999999.0) S->stp1 = step_init;
    }
    S->ftol = 1.0e-4;
    S->maxfev = 20;
L80:
    S->iter = S->iter + 1;
    S->info = 0;
    S->bound = S->iter - 1;
    if (S->iter == 1) goto L165;
    if (S->iter > m) S->bound = m;
    S->ys = 0.0;
    for (j = 1; j <= n; j++) S->ys = S->ys + w[S->iypt + S->npt + j - 1] * w[S->ispt + S->npt + j - 1];
    S->yy = 0.0;
    for (j = 1; j <= n; j++) S->yy = S->yy + w[S->iypt + S->npt + j - 1] * w[S->iypt + S->npt + j - 1];
    for (i = 0; i < n; i++) diag[i] = S->ys / S->yy;
L100:
    S->cp = S->point;
    if (S->point == 0) S->cp = m;
    w[n + S->cp - 1] = 1.0 / S->ys;
    for (i = 1; i <= n; i++) w[i - 1] = -g[i - 1];
    S->cp = S->point;
    for (i = 1; i <= S->bound; i++) {
        S->cp = S->cp - 1;
        if (S->cp == -1) S->cp = m - 1;
        sq = 0.0;
        for (j = 1; j <= n; j++) sq = sq + w[S->ispt + S->cp * n + j - 1] * w[j - 1];
        inmc = n + m + S->cp + 1;
        iycn = S->iypt + S->cp * n;
        w[inmc - 1] = w[n + S->cp] * sq;
        for (j = 1; j <= n; j++) w[j - 1] = w[j - 1] - w[inmc - 1] * w[iycn + j - 1];
    }
    for (i = 0; i < n; i++) w[i] = diag[i] * w[i];
    for (i = 1; i <= S->bound; i++) {
        yr = 0.0;
        for (j = 1; j <= n; j++) yr = yr + w[S->iypt + S->cp * n + j - 1] * w[j - 1];
        beta = w[n + S->cp] * yr;
        inmc = n + m + S->cp + 1;
        beta = w[inmc - 1] - beta;
        iscn = S->ispt + S->cp * n;
        for (j = 1; j <= n; j++) w[j - 1] = w[j - 1] + beta * w[iscn + j - 1];
        S->cp = S->cp + 1;
        if (S->cp == m) S->cp = 0;
    }
    for (i = 1; i <= n; i++) w[S->ispt + S->point * n + i - 1] = w[i - 1];
L165:
    S->nfev = 0;
    S->stp = 1.0;
    if (S->iter == 1) S->stp = S->stp1;
    for (i = 0; i < n; i++) w[i] = g[i];
L172:
    mcsrch(S, n, x, f, g, &w[S->ispt + S->point * n], &S->stp, S->ftol, xtol, S->maxfev, &S->info, &S->nfev,
           diag, c);
    if (S->info == -1) {
        *iflag = 1;
        return;
    }
    if (S->info != 1) {
        *iflag = -1;
        return;
    }
    S->nfun = S->nfun + S->nfev;
    *myinfo = S->info;
    if (c->boundt == 2 && S->info == 1) clip(c, x, n);
    S->npt = S->point * n;
    for (i = 1; i <= n; i++) {
        w[S->ispt + S->npt + i - 1] = S->stp * w[S->ispt + S->npt + i - 1];
        w[S->iypt + S->npt + i - 1] = g[i - 1] - w[i - 1];
    }
    S->point = S->point + 1;
    if (S->point == m) S->point = 0;
    S->gnorm = sqrt(orc_ddot(n, g, g));
    xnorm = sqrt(orc_ddot(n, x, x));
    xnorm = fmax(1.0, xnorm);
    if (S->gnorm / xnorm <= eps) S->finish = 1;
    if (S->finish) {
        *iflag = 0;
        return;
    }
    goto L80;
}

void orc_lbfgs(orc_lbfgs_t *S, int n, int m, double *x, double f, const double *g, double *diag, double eps,
               double xtol, double *w, int *iflag, int *myinfo, double xmin, double xmax, const float *mask,
               int boundt, int type, const double *xinit)
{
    orc_clip_t c = {0, xmin, xmax, mask, type, xinit, NULL, NULL, 0, boundt};
    lbfgs_core(S, n, m, x, f, g, diag, eps, xtol, w, iflag, myinfo, &c, 0, 999999.0);
}

/* LBFGSOC(N,M,X,F,G,.FALSE.,DIAG,IPRINT,EPS,XTOL,STEPINIT,W,IFLAG,MYINFO,...,
 * XMIN,XMAX,NPARAM,SPT1_OPT,BOUNDT) with the SAVE state in S instead of files. */
void orc_lbfgsoc(orc_lbfgs_t *S, int n, int m, double *x, double f, const double *g, double *diag, double eps,
                 double xtol, double step_init, double *w, int *iflag, int *myinfo, const double *xmin,
                 const double *xmax, int nparam, int stp1_opt, int boundt)
{
    orc_clip_t c = {1, 0.0, 0.0, NULL, 0, NULL, xmin, xmax, nparam, boundt};
    lbfgs_core(S, n, m, x, f, g, diag, eps, xtol, w, iflag, myinfo, &c, stp1_opt, step_init);
}

/*
 * Inversion.f90:29-189 l_bfgs: the in-process driver around LBFGS.
 * fct(ctx, x, g, f) evaluates the misfit f (double) and its gradient g (FP32,
 * the reference's sign convention: LBFGS receives gd = -g) at the FP32 model x.
 * On the first evaluation (eval == 0) g0 = maxval(x)/(100 maxval|g|) and
 * f0 = maxval(x)^2/|f| are fixed (FP32 and FP64 as in :117-121) and scale every
 * later g and f.  The LBFGS iterate is rounded through FP32 between calls
 * (x = sngl(xd); xd = dble(x)), as the reference does (:127-136).
 * Histories: f_eval[k] = unscaled misfit of evaluation k, accepted[k] = 1
 * where that call completed an iteration (MYINFO == 1); cap bounds both.
 * Returns the final IFLAG; *nevals, *niters count evaluations / iterations.
 */
typedef int (*orc_fctgdt)(void *ctx, const float *x, float *g, double *f);

int orc_l_bfgs(int n, float *x, int niter, int neval, int eval0, int msave, double xmin, double xmax,
               const float *mask, int const_type, int freeze_soft, const double *vpinit, orc_fctgdt fct,
               void *ctx, double *f_eval, int *accepted, int cap, int *nevals, int *niters)
{
    const double EPS = 1.0e-20, XTOL = 2.220446049250313e-16;
    double *wd = calloc((size_t)n * (2 * msave + 1) + 2 * msave, sizeof(double));
    double *diagd = calloc(n, sizeof(double)), *gd = calloc(n, sizeof(double));
    double *xd = calloc(n, sizeof(double)), *xsave = calloc(n, sizeof(double));
    double *xdsave = calloc(n, sizeof(double));
    float *g = calloc(n, sizeof(float));
    orc_lbfgs_t *S = orc_lbfgs_new();
    float g0 = 0.f;
    double f0 = 0.0, fd = 0.0;
    int i, iflag = 0, myinfo = 0, cont = 1, found = 0, count = 0, eval = eval0, k = 0;
    for (i = 0; i < n; i++) xsave[i] = (double)x[i];
    memcpy(xdsave, xsave, n * sizeof(double));
    while (cont && !found) {
        fct(ctx, x, g, &fd);
        if (k < cap) f_eval[k] = fd;
        if (eval == 0) {
            float mx = x[0], mg = fabsf(g[0]);
            for (i = 1; i < n; i++) {
                if (x[i] > mx) mx = x[i];
                if (fabsf(g[i]) > mg) mg = fabsf(g[i]);
            }
            g0 = mx / (100 * mg);
            f0 = (double)(mx * mx) / fabs(fd);
        }
        for (i = 0; i < n; i++) g[i] = g0 * g[i];
        fd = f0 * fd;
        for (i = 0; i < n; i++) {
            xd[i] = (double)x[i];
            gd[i] = -(double)g[i];
        }
        memcpy(xdsave, xd, n * sizeof(double));
        orc_lbfgs(S, n, msave, xd, fd, gd, diagd, EPS, XTOL, wd, &iflag, &myinfo, xmin, xmax, mask,
                  const_type, freeze_soft, vpinit);
        for (i = 0; i < n; i++) x[i] = (float)xd[i];
        if (k < cap) accepted[k] = myinfo == 1;
        if (myinfo == 1) {
            if (const_type == 2) orc_freeze_x(xmin, xmax, mask, freeze_soft, xdsave, n, vpinit);
            memcpy(xsave, xdsave, n * sizeof(double));
            myinfo = 0;
            count = count + 1;
            if (count >= niter) found = 1;
        }
        if (iflag <= 0) cont = 0;
        eval = eval + 1;
        k = k + 1;
        if (eval >= neval) cont = 0;
    }
    if (!cont)
        for (i = 0; i < n; i++) x[i] = (float)xsave[i];
    *nevals = k;
    *niters = count;
    free(wd); free(diagd); free(gd); free(xd); free(xsave); free(xdsave); free(g);
    orc_lbfgs_free(S);
    return iflag;
}
