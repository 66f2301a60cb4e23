// lbfgs.hip — device L-BFGS: restatement of libs/libinv/lbfgs.f
// (LBFGS :49-558, MCSRCH :652-982, MCSTEP :985-1210, FREEZE_X :1-47)
// with FP64 device vectors, and its C-ABI entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "../../include/seisfd.h"
#include "common.h"
#include "ftn_io.h"
#include "lbfgs_dev.h"
#include "lbfgs_host.h"

using namespace sfd;

static dim3 vgrid(long n) { return dim3((unsigned)std::min<long>((n + 255) / 256, 4096)); }

LbfgsDev::LbfgsDev(long n_, int m_, double xmin_, double xmax_, const float *mask_h, int boundt_, int type_,
                   const double *xinit_h)
    : n(n_), m(m_), xmin(xmin_), xmax(xmax_), boundt(boundt_), type(type_) {
    if (n < 1 || m < 1) throw EngineError(SFD_EINVAL, "L-BFGS needs n >= 1 and m >= 1");
    try {
        init(mask_h, xinit_h);
    } catch (...) {
        release();   // a constructor that throws runs no destructor
        throw;
    }
}

void LbfgsDev::init(const float *mask_h, const double *xinit_h) {
    HIPCHK(hipStreamCreate(&st));
    const size_t b = (size_t)n * sizeof(double);
    HIPCHK(hipMalloc(&x, b));
    HIPCHK(hipMalloc(&g, b));
    HIPCHK(hipMalloc(&q, b));
    HIPCHK(hipMalloc(&diag, b));
    HIPCHK(hipMalloc(&xinit, b));
    HIPCHK(hipMalloc(&S, b * m));
    HIPCHK(hipMalloc(&Y, b * m));
    HIPCHK(hipMalloc(&part, sizeof(double) * kDotBlocks));
    HIPCHK(hipMalloc(&res, sizeof(double)));
    HIPCHK(hipHostMalloc(&res_h, sizeof(double)));
    HIPCHK(hipMemsetAsync(S, 0, b * m, st));
    HIPCHK(hipMemsetAsync(Y, 0, b * m, st));
    if (xinit_h) HIPCHK(hipMemcpyAsync(xinit, xinit_h, b, hipMemcpyHostToDevice, st));
    else HIPCHK(hipMemsetAsync(xinit, 0, b, st));
    if (mask_h) {
        HIPCHK(hipMalloc(&mask, (size_t)n * sizeof(float)));
        HIPCHK(hipMemcpyAsync(mask, mask_h, (size_t)n * sizeof(float), hipMemcpyHostToDevice, st));
    }
    rho.assign(m, 0.0);
    alpha.assign(m, 0.0);
    HIPCHK(hipStreamSynchronize(st));
}

LbfgsDev::~LbfgsDev() { release(); }

void LbfgsDev::release() {
    for (void *p : {(void *)x, (void *)g, (void *)q, (void *)diag, (void *)xinit, (void *)S, (void *)Y,
                    (void *)part, (void *)res, (void *)mask})
        if (p) (void)hipFree(p);
    if (res_h) (void)hipHostFree(res_h);
    if (st) (void)hipStreamDestroy(st);
    x = g = q = diag = xinit = S = Y = part = res = res_h = nullptr;
    mask = nullptr;
    st = nullptr;
}

double LbfgsDev::dot(const double *a, const double *b) {
    hipLaunchKernelGGL(dot_partial_kernel, dim3(kDotBlocks), dim3(256), 0, st, a, b, n, part);
    hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(256), 0, st, (const double *)part, kDotBlocks, res);
    HIPCHK(hipMemcpyAsync(res_h, res, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return *res_h;
}
void LbfgsDev::axpby(double a, const double *xv, double b, double *y) {
    hipLaunchKernelGGL(axpby_kernel, vgrid(n), dim3(256), 0, st, n, a, xv, b, y);
}
void LbfgsDev::scale_copy(double a, const double *xv, double *y) {
    hipLaunchKernelGGL(copy_scaled_kernel, vgrid(n), dim3(256), 0, st, n, a, xv, y);
}
void LbfgsDev::freeze(double *xv) {
    if (nparam > 0) {   // FREEZE_XOC (nlinv_lbfgs.f:1-22): block j of n/nparam values; a remainder is left alone
        const long blk = n / nparam;
        for (int j = 0; j < nparam; j++)
            hipLaunchKernelGGL(clip_block_kernel, vgrid(blk), dim3(256), 0, st, (long)j * blk, blk, pxmin[j],
                               pxmax[j], xv);
        return;
    }
    hipLaunchKernelGGL(freeze_kernel, vgrid(n), dim3(256), 0, st, n, xmin, xmax, (const float *)mask, type, xv,
                       (const double *)xinit);
}

// MCSTEP (lbfgs.f:985-1210), scalar
static void mcstep(double *stx, double *fx, double *dx, double *sty, double *fy, double *dy, double *stp, double fp,
                   double dp, int *brackt, double stpmin, double stpmax, int *info) {
    int bound;
    double gamma, p, q, r, s, sgnd, stpc, stpf, stpq, theta;
    *info = 0;
    if ((*brackt && (*stp <= std::min(*stx, *sty) || *stp >= std::max(*stx, *sty))) || *dx * (*stp - *stx) >= 0.0 ||
        stpmax < stpmin)
        return;
    sgnd = dp * (*dx / fabs(*dx));
    if (fp > *fx) {
        *info = 1;
        bound = 1;
        theta = 3 * (*fx - fp) / (*stp - *stx) + *dx + dp;
        s = std::max(fabs(theta), std::max(fabs(*dx), fabs(dp)));
        gamma = s * sqrt((theta / s) * (theta / s) - (*dx / s) * (dp / s));
        if (*stp < *stx) gamma = -gamma;
        p = (gamma - *dx) + theta;
        q = ((gamma - *dx) + gamma) + dp;
        r = p / q;
        stpc = *stx + r * (*stp - *stx);
        stpq = *stx + ((*dx / ((*fx - fp) / (*stp - *stx) + *dx)) / 2) * (*stp - *stx);
        stpf = fabs(stpc - *stx) < fabs(stpq - *stx) ? stpc : stpc + (stpq - stpc) / 2;
        *brackt = 1;
    } else if (sgnd < 0.0) {
        *info = 2;
        bound = 0;
        theta = 3 * (*fx - fp) / (*stp - *stx) + *dx + dp;
        s = std::max(fabs(theta), std::max(fabs(*dx), fabs(dp)));
        gamma = s * sqrt((theta / s) * (theta / s) - (*dx / s) * (dp / s));
        if (*stp > *stx) gamma = -gamma;
        p = (gamma - dp) + theta;
        q = ((gamma - dp) + gamma) + *dx;
        r = p / q;
        stpc = *stp + r * (*stx - *stp);
        stpq = *stp + (dp / (dp - *dx)) * (*stx - *stp);
        stpf = fabs(stpc - *stp) > fabs(stpq - *stp) ? stpc : stpq;
        *brackt = 1;
    } else if (fabs(dp) < fabs(*dx)) {
        *info = 3;
        bound = 1;
        theta = 3 * (*fx - fp) / (*stp - *stx) + *dx + dp;
        s = std::max(fabs(theta), std::max(fabs(*dx), fabs(dp)));
        gamma = s * sqrt(std::max(0.0, (theta / s) * (theta / s) - (*dx / s) * (dp / s)));
        if (*stp > *stx) gamma = -gamma;
        p = (gamma - dp) + theta;
        q = (gamma + (*dx - dp)) + gamma;
        r = p / q;
        if (r < 0.0 && gamma != 0.0) stpc = *stp + r * (*stx - *stp);
        else if (*stp > *stx) stpc = stpmax;
        else stpc = stpmin;
        stpq = *stp + (dp / (dp - *dx)) * (*stx - *stp);
        if (*brackt) stpf = fabs(*stp - stpc) < fabs(*stp - stpq) ? stpc : stpq;
        else stpf = fabs(*stp - stpc) > fabs(*stp - stpq) ? stpc : stpq;
    } else {
        *info = 4;
        bound = 0;
        if (*brackt) {
            theta = 3 * (fp - *fy) / (*sty - *stp) + *dy + dp;
            s = std::max(fabs(theta), std::max(fabs(*dy), fabs(dp)));
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
        *sty = *stp; *fy = fp; *dy = dp;
    } else {
        if (sgnd < 0.0) { *sty = *stx; *fy = *fx; *dy = *dx; }
        *stx = *stp; *fx = fp; *dx = dp;
    }
    stpf = std::min(stpmax, stpf);
    stpf = std::max(stpmin, stpf);
    *stp = stpf;
    if (*brackt && bound) {
        // 0.66 is a default-REAL literal in lbfgs.f:1203-1206: its FP32 value, widened
        const double p66r = (double)0.66f;
        if (*sty > *stx) *stp = std::min(*stx + p66r * (*sty - *stx), *stp);
        else *stp = std::max(*stx + p66r * (*sty - *stx), *stp);
    }
}

// MCSRCH (lbfgs.f:652-982); s = search direction (device), WA = diag.
void LbfgsDev::mcsrch(double f, double *s) {
    const double p5 = 0.5, p66 = 0.66, zero = 0.0;
    const double xtol = xtol_;
    if (info == -1) goto L45;
    infoc = 1;
    if (n <= 0 || stp <= zero || ftol < zero || gtol < zero || xtol < zero || stpmin < zero || stpmax < stpmin ||
        maxfev <= 0)
        return;
    dginit = dot(g, s);
    if (dginit >= zero) return;   // not a descent direction (lbfgs.f:810-814)
    brackt = 0;
    stage1 = 1;
    nfev = 0;
    finit = f;
    dgtest = ftol * dginit;
    width = stpmax - stpmin;
    width1 = width / p5;
    scale_copy(1.0, x, diag);
    stx = zero; fx = finit; dgx = dginit;
    sty = zero; fy = finit; dgy = dginit;
L30:
    if (brackt) {
        stmin = std::min(stx, sty);
        stmax = std::max(stx, sty);
    } else {
        stmin = stx;
        stmax = stp + xtrapf * (stp - stx);
    }
    stp = std::max(stp, stpmin);
    stp = std::min(stp, stpmax);
    if ((brackt && (stp <= stmin || stp >= stmax)) || nfev >= maxfev - 1 || infoc == 0 ||
        (brackt && stmax - stmin <= xtol * stmax))
        stp = stx;
    // X = WA + STP*S
    scale_copy(1.0, diag, x);
    axpby(stp, s, 1.0, x);
    if (boundt == 1) freeze(x);
    info = -1;
    if (on_mcsrch_write) on_mcsrch_write();
    return;
L45:
    info = 0;
    nfev = nfev + 1;
    {
        dg = dot(g, s);
        ftest1 = finit + stp * dgtest;
        if ((brackt && (stp <= stmin || stp >= stmax)) || infoc == 0) info = 6;
        if (stp == stpmax && f <= ftest1 && dg <= dgtest) info = 5;
        if (stp == stpmin && (f > ftest1 || dg >= dgtest)) info = 4;
        if (nfev >= maxfev) info = 3;
        if (brackt && stmax - stmin <= xtol * stmax) info = 2;
        if (f <= ftest1 && fabs(dg) <= gtol * (-dginit)) info = 1;
        if (info != 0) {
            if (on_mcsrch_write) on_mcsrch_write();
            return;
        }
        if (stage1 && f <= ftest1 && dg >= std::min(ftol, gtol) * dginit) stage1 = 0;
        if (stage1 && f <= fx && f > ftest1) {
            fm = f - stp * dgtest;
            fxm = fx - stx * dgtest;
            fym = fy - sty * dgtest;
            dgm = dg - dgtest;
            dgxm = dgx - dgtest;
            dgym = dgy - dgtest;
            mcstep(&stx, &fxm, &dgxm, &sty, &fym, &dgym, &stp, fm, dgm, &brackt, stmin, stmax, &infoc);
            fx = fxm + stx * dgtest;
            fy = fym + sty * dgtest;
            dgx = dgxm + dgtest;
            dgy = dgym + dgtest;
        } else {
            mcstep(&stx, &fx, &dgx, &sty, &fy, &dgy, &stp, f, dg, &brackt, stmin, stmax, &infoc);
        }
        if (brackt) {
            if (fabs(sty - stx) >= p66 * width1) stp = stx + p5 * (sty - stx);
            width1 = width;
            width = fabs(sty - stx);
        }
    }
    goto L30;
}

// The two-loop recursion of lbfgs.f:367-420 (DIAGCO = .false.): S(point) = -H g,
// with H0 = (ys/yy) I, or the learned H0 when use_h0.
void LbfgsDev::two_loop(bool use_h0) {
    used_h0 = use_h0;
    ys = dot(Yv(npt), Sv(npt));
    yy = dot(Yv(npt), Yv(npt));
    hipLaunchKernelGGL(fill_kernel, vgrid(n), dim3(256), 0, st, n, ys / yy, diag);
    cp = point == 0 ? m : point;
    rho[cp - 1] = 1.0 / ys;
    scale_copy(-1.0, g, q);
    cp = point;
    for (int i = 1; i <= bound; i++) {
        cp = cp - 1;
        if (cp == -1) cp = m - 1;
        sq = dot(Sv(cp), q);
        alpha[cp] = rho[cp] * sq;
        axpby(-alpha[cp], Yv(cp), 1.0, q);
        inmc = (int)(n + m + cp + 1);                 // the reference's W indexes, 1-based
        iycn = (int)(n + 2 * m + (long)m * n + (long)cp * n);
    }
    if (use_h0) h0(h0_ctx, q, n, st);     // learned initial inverse Hessian
    else hipLaunchKernelGGL(mul_kernel, vgrid(n), dim3(256), 0, st, n, (const double *)diag, q);
    for (int i = 1; i <= bound; i++) {
        yr = dot(Yv(cp), q);
        beta = rho[cp] * yr;
        beta = alpha[cp] - beta;
        axpby(beta, Sv(cp), 1.0, q);
        inmc = (int)(n + m + cp + 1);
        iscn = (int)(n + 2 * m + (long)cp * n);
        cp = cp + 1;
        if (cp == m) cp = 0;
    }
    scale_copy(1.0, q, Sv(point));
}

void LbfgsDev::mul_diag(double *v) {
    hipLaunchKernelGGL(mul_kernel, vgrid(n), dim3(256), 0, st, n, (const double *)diag, v);
}

// Give up the current (preconditioned) line search and restart it from its
// starting point with the plain direction: X = WA, G = W(1:N), F = FINIT.
// Returns the IFLAG of the restarted search (1: evaluate at the new X).
int LbfgsDev::restart_plain() {
    scale_copy(1.0, diag, x);
    scale_copy(1.0, q, g);
    const double f0 = finit;
    two_loop(false);
    nfev = 0;
    stp = 1.0;
    scale_copy(1.0, g, q);
    info = 0;
    mcsrch(f0, Sv(point));
    return info == -1 ? 1 : -1;
}

// LBFGS (lbfgs.f:49-558), DIAGCO = .false.
void LbfgsDev::step(double f, int *iflag, int *myinfo, double eps, double xtol) {
    xtol_ = xtol;
    if (*iflag == 0) goto L10;
    if (*iflag == 1) goto L172;
L10:
    iter = 0;
    if (n <= 0 || m <= 0) {
        *iflag = -3;
        return;
    }
    if (gtol <= 1.0e-04) gtol = 9.0e-01;
    nfun = 1;
    point = 0;
    finish = 0;
    hipLaunchKernelGGL(fill_kernel, vgrid(n), dim3(256), 0, st, n, 1.0, diag);
    scale_copy(-1.0, g, Sv(0));                  // W(ISPT+I) = -G(I)*DIAG(I), DIAG = 1
    gnorm = sqrt(dot(g, g));
    xnorm2 = sqrt(dot(x, x));
    if (stp1_opt < 0) {
        stp1 = boundt == 0 ? 1.0 / gnorm : xnorm2 / gnorm / 100;     // lbfgs.f:333-338
    } else {
        stp1 = stp1_opt == 1 ? xnorm2 / gnorm / 100 : 1.0 / gnorm;   // nlinv_lbfgs.f:336-341
        if (step_init != 999999.0) stp1 = step_init;                  // :342-344
    }
    ftol = 1.0e-4;
    maxfev = 20;
L80:
    iter = iter + 1;
    info = 0;
    used_h0 = false;
    bound = iter - 1;
    if (iter == 1) goto L165;
    if (iter > m) bound = m;
    two_loop(h0 && h0_active);
    if (used_h0 && dot(g, Sv(point)) >= 0.0) {   // the learned H0 gave no descent direction
        h0_rejects++;
        two_loop(false);
    }
L165:
    nfev = 0;
    stp = 1.0;
    if (iter == 1) stp = stp1;
    scale_copy(1.0, g, q);                       // W(I) = G(I): gradient kept for Y
L172:
    mcsrch(f, Sv(point));
    if (info == -1) {
        *iflag = 1;
        if (on_lbfgs_write) on_lbfgs_write();
        return;
    }
    if (info != 1) {
        *iflag = -1;
        return;
    }
    nfun = nfun + nfev;
    *myinfo = info;
    if (boundt == 2 && info == 1) freeze(x);
    npt = point;                                  // (the reference's NPT = POINT*N)
    scale_copy(stp, Sv(npt), Sv(npt));
    hipLaunchKernelGGL(sub_kernel, vgrid(n), dim3(256), 0, st, n, (const double *)g, (const double *)q, Yv(npt));
    // accepted pair: gradient at the start of the line search (q) and the step taken
    if (on_pair) on_pair(pair_ctx, q, Sv(npt), Yv(npt), n, st);
    point = point + 1;
    if (point == m) point = 0;
    gnorm = sqrt(dot(g, g));
    xnorm = sqrt(dot(x, x));
    xnorm = std::max(1.0, xnorm);
    if (gnorm / xnorm <= eps) finish = 1;
    if (finish) {
        *iflag = 0;
        if (on_lbfgs_write) on_lbfgs_write();
        return;
    }
    if (on_lbfgs_write) on_lbfgs_write();
    goto L80;
}

// ---------------------------------------------------------------------------
// Optimiser state on file (the out-of-core protocol of nlinv_lbfgs.f, which
// keeps the LBFGSOC/MCSRCHOC SAVE scalars in LbfgsDat.dat / McsrchDat.dat and
// W, DIAG in WorkArray.dat / Diag.dat): one binary file, scalars exact, ended
// by a 64-bit FNV-1a checksum of everything before it (8-byte words, then any
// tail bytes), so a damaged file is refused rather than resumed from.
namespace {
constexpr char kStateMagic[8] = {'S', 'F', 'D', 'L', 'B', 'O', 'C', '2'};
struct StateHead {
    char magic[8];
    int64_t n;
    int32_t m, nparam;
    int32_t iter, nfun, point, maxfev, info, bound, npt, cp, nfev, finish, infoc, brackt, stage1, pad;
    double gnorm, xnorm2, stp1, ftol, stp, ys, yy, gtol, stpmin, stpmax, xtrapf;
    double dginit, dgtest, finit, width, width1, stx, fx, dgx, sty, fy, dgy, stmin, stmax;
};

struct File {
    FILE *f = nullptr;
    uint64_t sum = 14695981039346656037ull;   // FNV-1a offset basis
    ~File() { if (f) fclose(f); }
    void add(const void *p, size_t b) {
        const unsigned char *c = (const unsigned char *)p;
        size_t k = 0;
        for (; k + 8 <= b; k += 8) {
            uint64_t w;
            memcpy(&w, c + k, 8);
            sum = (sum ^ w) * 1099511628211ull;
        }
        for (; k < b; k++) sum = (sum ^ c[k]) * 1099511628211ull;
    }
};

void wr(File &fo, const void *p, size_t b) {
    if (b && fwrite(p, 1, b, fo.f) != b) throw EngineError(SFD_EIO, "L-BFGS state: write failed");
    fo.add(p, b);
}
void rd(File &fi, void *p, size_t b) {
    if (b && fread(p, 1, b, fi.f) != b) throw EngineError(SFD_EIO, "L-BFGS state: file truncated");
    fi.add(p, b);
}
}  // namespace

void LbfgsDev::save_state(const char *path) {
    StateHead h{};
    memcpy(h.magic, kStateMagic, 8);
    h.n = n; h.m = m; h.nparam = nparam;
    h.iter = iter; h.nfun = nfun; h.point = point; h.maxfev = maxfev; h.info = info; h.bound = bound;
    h.npt = npt; h.cp = cp; h.nfev = nfev; h.finish = finish; h.infoc = infoc; h.brackt = brackt;
    h.stage1 = stage1;
    h.gnorm = gnorm; h.xnorm2 = xnorm2; h.stp1 = stp1; h.ftol = ftol; h.stp = stp; h.ys = ys; h.yy = yy;
    h.gtol = gtol; h.stpmin = stpmin; h.stpmax = stpmax; h.xtrapf = xtrapf;
    h.dginit = dginit; h.dgtest = dgtest; h.finit = finit; h.width = width; h.width1 = width1; h.stx = stx;
    h.fx = fx; h.dgx = dgx; h.sty = sty; h.fy = fy; h.dgy = dgy; h.stmin = stmin; h.stmax = stmax;
    const std::string tmp = std::string(path) + ".tmp";
    {
        File fo;
        fo.f = fopen(tmp.c_str(), "wb");
        if (!fo.f) throw EngineError(SFD_EIO, "L-BFGS state: cannot create " + tmp);
        wr(fo, &h, sizeof h);
        wr(fo, pxmin.data(), sizeof(double) * nparam);
        wr(fo, pxmax.data(), sizeof(double) * nparam);
        wr(fo, rho.data(), sizeof(double) * m);
        wr(fo, alpha.data(), sizeof(double) * m);
        std::vector<double> buf(n);
        auto put = [&](const double *dv) {
            HIPCHK(hipMemcpyAsync(buf.data(), dv, sizeof(double) * n, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            wr(fo, buf.data(), sizeof(double) * n);
        };
        put(q);
        put(diag);
        for (int k = 0; k < m; k++) put(Sv(k));
        for (int k = 0; k < m; k++) put(Yv(k));
        const uint64_t sum = fo.sum;
        if (fwrite(&sum, 1, 8, fo.f) != 8 || fflush(fo.f) != 0) throw EngineError(SFD_EIO, "L-BFGS state: write failed");
    }
    if (rename(tmp.c_str(), path) != 0) throw EngineError(SFD_EIO, std::string("L-BFGS state: cannot rename to ") + path);
}

void LbfgsDev::load_state(const char *path) {
    File fi;
    fi.f = fopen(path, "rb");
    if (!fi.f) throw EngineError(SFD_EIO, std::string("L-BFGS state: cannot open ") + path);
    StateHead h;
    rd(fi, &h, sizeof h);
    if (memcmp(h.magic, kStateMagic, 8) != 0) throw EngineError(SFD_EIO, "L-BFGS state: not a state file");
    if (h.n != n || h.m != m || h.nparam != nparam)
        throw EngineError(SFD_EIO, "L-BFGS state: n, m or nparam differ from the run's");
    std::vector<double> b0(nparam), b1(nparam);
    rd(fi, b0.data(), sizeof(double) * nparam);
    rd(fi, b1.data(), sizeof(double) * nparam);
    if (b0 != pxmin || b1 != pxmax) throw EngineError(SFD_EIO, "L-BFGS state: bounds differ from the run's");
    iter = h.iter; nfun = h.nfun; point = h.point; maxfev = h.maxfev; info = h.info; bound = h.bound;
    npt = h.npt; cp = h.cp; nfev = h.nfev; finish = h.finish; infoc = h.infoc; brackt = h.brackt;
    stage1 = h.stage1;
    gnorm = h.gnorm; xnorm2 = h.xnorm2; stp1 = h.stp1; ftol = h.ftol; stp = h.stp; ys = h.ys; yy = h.yy;
    gtol = h.gtol; stpmin = h.stpmin; stpmax = h.stpmax; xtrapf = h.xtrapf;
    dginit = h.dginit; dgtest = h.dgtest; finit = h.finit; width = h.width; width1 = h.width1; stx = h.stx;
    fx = h.fx; dgx = h.dgx; sty = h.sty; fy = h.fy; dgy = h.dgy; stmin = h.stmin; stmax = h.stmax;
    rd(fi, rho.data(), sizeof(double) * m);
    rd(fi, alpha.data(), sizeof(double) * m);
    std::vector<double> buf(n);
    auto get = [&](double *dv) {
        rd(fi, buf.data(), sizeof(double) * n);
        HIPCHK(hipMemcpyAsync(dv, buf.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
    };
    get(q);
    get(diag);
    for (int k = 0; k < m; k++) get(Sv(k));
    for (int k = 0; k < m; k++) get(Yv(k));
    const uint64_t sum = fi.sum;
    uint64_t stored = 0;
    if (fread(&stored, 1, 8, fi.f) != 8) throw EngineError(SFD_EIO, "L-BFGS state: file truncated");
    if (stored != sum) throw EngineError(SFD_EIO, "L-BFGS state: checksum mismatch (damaged file)");
    if (fgetc(fi.f) != EOF) throw EngineError(SFD_EIO, "L-BFGS state: trailing data");
}

// ---------------------------------------------------------------------------
// Inversion.f90:29-189 l_bfgs around the device LBFGS.  The control flow
// (first-evaluation scaling g0/f0, FP32 round trip of the iterate, FREEZE_X
// of the accepted model, niter/neval stops, the final model choice) follows
// the reference line by line; the vector work is on the device in FP64.
__global__ void lbfgs_prep_kernel(long n, const float *__restrict__ xf, float *__restrict__ gf, float g0,
                                  double *__restrict__ xd, double *__restrict__ gd) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float gs = g0 * gf[i];
        gf[i] = gs;
        xd[i] = (double)xf[i];
        gd[i] = -(double)gs;
    }
}
__global__ void to_float_kernel(long n, const double *__restrict__ xd, float *__restrict__ xf) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        xf[i] = (float)xd[i];
}

int lbfgs_drive(const sfd_lbfgs_opts &o, long n, float *x, const float *mask, const double *vpinit,
                const LbfgsFct &fct, sfd_lbfgs_report *rep, LbfgsHooks *hooks) {
    const double EPS = 1.0e-20, XTOL = 2.220446049250313e-16;   // Inversion.f90:79-80
    const int msave = o.m > 0 ? o.m : 5;
    std::vector<float> ones;
    if (!mask) {
        ones.assign(n, 1.0f);
        mask = ones.data();
    }
    std::vector<double> vpi;
    if (!vpinit) {
        vpi.assign(x, x + n);
        vpinit = vpi.data();
    }
    LbfgsDev d(n, msave, o.xmin, o.xmax, mask, o.const_type, o.freeze_soft, vpinit);
    if (hooks) hooks->attach(d);
    float *xf = nullptr, *gf = nullptr;
    double *xdsave = nullptr, *xsave = nullptr;
    HIPCHK(hipMalloc(&xf, n * sizeof(float)));
    HIPCHK(hipMalloc(&gf, n * sizeof(float)));
    HIPCHK(hipMalloc(&xdsave, n * sizeof(double)));
    HIPCHK(hipMalloc(&xsave, n * sizeof(double)));
    struct Free {
        void *p[4];
        ~Free() { for (void *q : p) if (q) (void)hipFree(q); }
    } fr{{xf, gf, xdsave, xsave}};
    const dim3 vg = vgrid(n), vb(256);
    std::vector<float> g(n);
    float g0 = 0.f;
    double f0 = 0.0, fd = 0.0;
    int iflag = 0, myinfo = 0, cont = 1, found = 0, count = 0, eval = o.eval0, k = 0;
    const auto t0 = std::chrono::steady_clock::now();
    auto secs = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
    HIPCHK(hipMemcpyAsync(xf, x, n * sizeof(float), hipMemcpyHostToDevice, d.st));
    hipLaunchKernelGGL(lbfgs_prep_kernel, vg, vb, 0, d.st, n, (const float *)xf, gf, 0.f, xsave, d.g);
    HIPCHK(hipMemcpyAsync(xdsave, xsave, n * sizeof(double), hipMemcpyDeviceToDevice, d.st));
    HIPCHK(hipStreamSynchronize(d.st));
    while (cont && !found) {
        int rc;
        try {
            rc = fct(x, g.data(), &fd);
        } catch (const EngineError &err) {
            // A trial model along a preconditioner's direction that the
            // evaluation rejects (non-finite or non-positive velocity, Courant
            // limit, the sampling stop of a too-low velocity, a non-finite
            // wavefield) is a failed preconditioned search: the hook restarts
            // it plain from the same model.  Any other failure propagates.
            const bool model_range = err.code == SFD_EINVAL || err.code == SFD_EUNSTABLE ||
                                     err.code == SFD_EALIASED || err.code == SFD_ENOTFINITE;
            if (!(hooks && d.used_h0 && model_range)) throw;
            iflag = hooks->rejected_trial(d);
            if (iflag != 1) throw;
            hipLaunchKernelGGL(to_float_kernel, vg, vb, 0, d.st, n, (const double *)d.x, xf);
            HIPCHK(hipMemcpyAsync(x, xf, n * sizeof(float), hipMemcpyDeviceToHost, d.st));
            HIPCHK(hipStreamSynchronize(d.st));
            if (rep && k < rep->hist_cap) {
                if (rep->f_eval) rep->f_eval[k] = std::nan("");
                if (rep->t_eval) rep->t_eval[k] = secs();
                if (rep->accepted) rep->accepted[k] = 0;
            }
            eval = eval + 1;
            k = k + 1;
            if (eval >= o.neval) cont = 0;
            continue;
        }
        if (rc != 0) throw EngineError(rc < 0 ? rc : SFD_ESTATE, "misfit/gradient evaluation failed");
        if (rep && k < rep->hist_cap) {
            if (rep->f_eval) rep->f_eval[k] = fd;
            if (rep->t_eval) rep->t_eval[k] = secs();
        }
        if (eval == 0) {
            float mx = x[0], mg = std::fabs(g[0]);
            for (long i = 1; i < n; i++) {
                mx = std::max(mx, x[i]);
                mg = std::max(mg, std::fabs(g[i]));
            }
            g0 = mx / (100 * mg);
            f0 = (double)(mx * mx) / std::fabs(fd);
        }
        fd = f0 * fd;
        HIPCHK(hipMemcpyAsync(xf, x, n * sizeof(float), hipMemcpyHostToDevice, d.st));
        HIPCHK(hipMemcpyAsync(gf, g.data(), n * sizeof(float), hipMemcpyHostToDevice, d.st));
        hipLaunchKernelGGL(lbfgs_prep_kernel, vg, vb, 0, d.st, n, (const float *)xf, gf, g0, d.x, d.g);
        HIPCHK(hipMemcpyAsync(xdsave, d.x, n * sizeof(double), hipMemcpyDeviceToDevice, d.st));
        if (hooks) hooks->before_step(d, fd);
        d.step(fd, &iflag, &myinfo, EPS, XTOL);
        if (hooks) iflag = hooks->after_step(d, iflag);
        hipLaunchKernelGGL(to_float_kernel, vg, vb, 0, d.st, n, (const double *)d.x, xf);
        HIPCHK(hipMemcpyAsync(x, xf, n * sizeof(float), hipMemcpyDeviceToHost, d.st));
        HIPCHK(hipStreamSynchronize(d.st));
        if (rep && k < rep->hist_cap && rep->accepted) rep->accepted[k] = myinfo == 1;
        if (myinfo == 1) {
            if (o.const_type == 2) d.freeze(xdsave);
            HIPCHK(hipMemcpyAsync(xsave, xdsave, n * sizeof(double), hipMemcpyDeviceToDevice, d.st));
            myinfo = 0;
            count = count + 1;
            if (count >= o.niter) found = 1;
            if (hooks) hooks->accepted(d, count);
        }
        if (iflag <= 0) cont = 0;
        eval = eval + 1;
        k = k + 1;
        if (eval >= o.neval) cont = 0;
    }
    if (!cont) {
        hipLaunchKernelGGL(to_float_kernel, vg, vb, 0, d.st, n, (const double *)xsave, xf);
        HIPCHK(hipMemcpyAsync(x, xf, n * sizeof(float), hipMemcpyDeviceToHost, d.st));
    }
    HIPCHK(hipStreamSynchronize(d.st));
    if (rep) {
        rep->iterations = count;
        rep->evaluations = k;
        rep->iflag = iflag;
        rep->info = d.info;
        if (hooks) hooks->report(rep);
        rep->seconds = secs();
        rep->nhist = std::min(k, rep->hist_cap);
    }
    return iflag;
}

// ---------------------------------------------------------------------------
// The state files of progs/inversion/L_BFGS.f90 with lbfgs_type 1: LBFGSOC's
// and MCSRCHOC's SAVE scalars in the list-directed text files LbfgsDat.dat and
// McsrchDat.dat (nlinv_lbfgs.f:276-281, :755-760), W and DIAG as one
// unformatted record each (WorkArray.dat, Diag.dat; L_BFGS.f90:69-78 and
// nlinv_read_array_params.f90:237-248), IFLAG and MYINFO as one integer each
// (Iflag.dat, Info.dat).  Reads take the input names, writes the *Out names
// (nlinv_read_array_params.f90:139-151).  The record order of the two text
// files and of W (q, rho, alpha, the m steps S, the m gradient changes Y) is
// the reference's, so either program can resume from the other's files.
namespace lbx {

void read_lbfgs_dat(LbfgsDev &d, const std::string &path, bool check_layout) {
    ftn::ListReader r(path);
    double one = 1.0, zero = 0.0;
    int ispt = 0, iypt = 0, npt_ref = 0, i_ = 0;
    r.statement();
    r.real(&d.gtol); r.real(&one); r.real(&zero); r.real(&d.gnorm); r.real(&d.stp1); r.real(&d.ftol);
    r.real(&d.stpmin);
    r.statement();
    r.real(&d.stpmax); r.real(&d.stp); r.real(&d.ys); r.real(&d.yy); r.real(&d.sq); r.real(&d.yr);
    r.real(&d.beta); r.real(&d.xnorm);
    r.statement();
    r.integer(&d.mp); r.integer(&d.lp); r.integer(&d.iter); r.integer(&d.nfun); r.integer(&d.point);
    r.integer(&ispt); r.integer(&iypt); r.integer(&d.maxfev); r.integer(&d.info);
    r.statement();
    r.integer(&d.bound); r.integer(&npt_ref); r.integer(&d.cp); r.integer(&i_); r.integer(&d.nfev);
    r.integer(&d.inmc); r.integer(&d.iycn); r.integer(&d.iscn);
    r.statement();
    r.logical(&d.finish);
    // ONE and ZERO are variables of LBFGSOC that the file overwrites
    if (one != 1.0 || zero != 0.0) throw EngineError(SFD_EIO, path + ": ONE and ZERO are not 1 and 0");
    if (!check_layout) return;
    const long n = d.n, m = d.m;
    if (ispt != n + 2 * m || iypt != n + 2 * m + n * m)
        throw EngineError(SFD_EIO, path + ": ISPT/IYPT were written for another N or M");
    if (npt_ref % n != 0 || npt_ref / n >= m || d.point < 0 || d.point >= m || d.cp < 0 || d.cp > m)
        throw EngineError(SFD_EIO, path + ": NPT, POINT or CP out of range");
    d.npt = (int)(npt_ref / n);   // the reference's NPT is the offset POINT*N
}

void write_lbfgs_dat(const LbfgsDev &d, const std::string &path) {
    ftn::ListWriter w;
    const long n = d.n, m = d.m;
    w.real(d.gtol); w.real(1.0); w.real(0.0); w.real(d.gnorm); w.real(d.stp1); w.real(d.ftol); w.real(d.stpmin);
    w.end_record();
    w.real(d.stpmax); w.real(d.stp); w.real(d.ys); w.real(d.yy); w.real(d.sq); w.real(d.yr); w.real(d.beta);
    w.real(d.xnorm);
    w.end_record();
    w.integer(d.mp); w.integer(d.lp); w.integer(d.iter); w.integer(d.nfun); w.integer(d.point);
    w.integer((int)(n + 2 * m)); w.integer((int)(n + 2 * m + n * m)); w.integer(d.maxfev); w.integer(d.info);
    w.end_record();
    // I is LBFGSOC's loop index, N + 1 after the loops that precede every write
    w.integer(d.bound); w.integer((int)(d.npt * n)); w.integer(d.cp); w.integer((int)(n + 1)); w.integer(d.nfev);
    w.integer(d.inmc); w.integer(d.iycn); w.integer(d.iscn);
    w.end_record();
    w.logical(d.finish);
    w.end_record();
    w.save(path);
}

void read_mcsrch_dat(LbfgsDev &d, const std::string &path) {
    ftn::ListReader r(path);
    int j = 0;
    r.statement();
    r.integer(&d.infoc); r.integer(&j);
    r.statement();
    r.logical(&d.brackt); r.logical(&d.stage1);
    r.statement();
    r.real(&d.dg); r.real(&d.dgm); r.real(&d.dginit); r.real(&d.dgtest); r.real(&d.dgx); r.real(&d.dgxm);
    r.real(&d.dgy); r.real(&d.dgym);
    r.statement();
    r.real(&d.finit); r.real(&d.ftest1); r.real(&d.fm); r.real(&d.fx); r.real(&d.fxm); r.real(&d.fy); r.real(&d.fym);
    r.real(&d.stx); r.real(&d.sty);
    r.statement();
    r.real(&d.stmin); r.real(&d.stmax); r.real(&d.width); r.real(&d.width1);
}

void write_mcsrch_dat(const LbfgsDev &d, const std::string &path) {
    ftn::ListWriter w;
    // J is MCSRCHOC's loop index, N + 1 after the loops that precede every write
    w.integer(d.infoc); w.integer((int)(d.n + 1));
    w.end_record();
    w.logical(d.brackt); w.logical(d.stage1);
    w.end_record();
    w.real(d.dg); w.real(d.dgm); w.real(d.dginit); w.real(d.dgtest); w.real(d.dgx); w.real(d.dgxm); w.real(d.dgy);
    w.real(d.dgym);
    w.end_record();
    w.real(d.finit); w.real(d.ftest1); w.real(d.fm); w.real(d.fx); w.real(d.fxm); w.real(d.fy); w.real(d.fym);
    w.real(d.stx); w.real(d.sty);
    w.end_record();
    w.real(d.stmin); w.real(d.stmax); w.real(d.width); w.real(d.width1);
    w.end_record();
    w.save(path);
}

// W(1:N) = q, W(N+1:N+M) = rho, W(N+M+1:N+2M) = alpha, then S and Y
// (nlinv_lbfgs.f:308-322), staged through one host vector of n values.
void read_work(LbfgsDev &d, const std::string &path) {
    const size_t n = d.n, m = d.m;
    ftn::RecordReader r(path, n * (2 * m + 1) + 2 * m);
    std::vector<double> buf(n);
    auto get = [&](double *dv) {
        r.read(buf.data(), n);
        HIPCHK(hipMemcpyAsync(dv, buf.data(), sizeof(double) * n, hipMemcpyHostToDevice, d.st));
        HIPCHK(hipStreamSynchronize(d.st));
    };
    get(d.q);
    r.read(d.rho.data(), m);
    r.read(d.alpha.data(), m);
    for (size_t k = 0; k < m; k++) get(d.Sv((int)k));
    for (size_t k = 0; k < m; k++) get(d.Yv((int)k));
}

void write_work(const LbfgsDev &d, const std::string &path) {
    const size_t n = d.n, m = d.m;
    ftn::RecordWriter w(path, n * (2 * m + 1) + 2 * m);
    std::vector<double> buf(n);
    auto put = [&](const double *dv) {
        HIPCHK(hipMemcpyAsync(buf.data(), dv, sizeof(double) * n, hipMemcpyDeviceToHost, d.st));
        HIPCHK(hipStreamSynchronize(d.st));
        w.write(buf.data(), n);
    };
    put(d.q);
    w.write(d.rho.data(), m);
    w.write(d.alpha.data(), m);
    for (size_t k = 0; k < m; k++) put(d.Sv((int)k));
    for (size_t k = 0; k < m; k++) put(d.Yv((int)k));
    w.close();
}

void read_vector(LbfgsDev &d, double *dv, const std::string &path) {
    ftn::RecordReader r(path, d.n);
    std::vector<double> buf(d.n);
    r.read(buf.data(), d.n);
    HIPCHK(hipMemcpyAsync(dv, buf.data(), sizeof(double) * d.n, hipMemcpyHostToDevice, d.st));
    HIPCHK(hipStreamSynchronize(d.st));
}

void write_vector(const LbfgsDev &d, const double *dv, const std::string &path) {
    ftn::RecordWriter w(path, d.n);
    std::vector<double> buf(d.n);
    HIPCHK(hipMemcpyAsync(buf.data(), dv, sizeof(double) * d.n, hipMemcpyDeviceToHost, d.st));
    HIPCHK(hipStreamSynchronize(d.st));
    w.write(buf.data(), d.n);
    w.close();
}

int read_flag(const std::string &path) {   // absent: 0 (nlinv_read_array_params.f90:154-174)
    int v = 0;
    if (!ftn::exists(path)) return 0;
    ftn::ListReader r(path);
    r.statement();
    r.integer(&v);
    return v;
}

}  // namespace lbx

// ---------------------------------------------------------------------------
struct sfd_lbfgs {
    LbfgsDev *d = nullptr;
    std::string err;
    ~sfd_lbfgs() { delete d; }
};

extern "C" {

int sfd_lbfgs_create(sfd_lbfgs_handle *lb, int n, int m, double xmin, double xmax, const float *mask, int boundt,
                     int freeze_soft, const double *xinit) {
    if (!lb) return SFD_EINVAL;
    *lb = nullptr;
    return abi_call(nullptr, [&]() {
        std::unique_ptr<sfd_lbfgs> h(new sfd_lbfgs());
        h->d = new LbfgsDev(n, m, xmin, xmax, mask, boundt, freeze_soft, xinit);
        *lb = h.release();
    });
}

int sfd_lbfgs_step(sfd_lbfgs_handle lb, double *x, double f, const double *g, int *iflag, int *info) {
    if (!lb || !lb->d || !x || !g || !iflag || !info) return SFD_EINVAL;
    LbfgsDev &d = *lb->d;
    return abi_call(&lb->err, [&]() {
        const size_t b = (size_t)d.n * sizeof(double);
        HIPCHK(hipMemcpyAsync(d.x, x, b, hipMemcpyHostToDevice, d.st));
        HIPCHK(hipMemcpyAsync(d.g, g, b, hipMemcpyHostToDevice, d.st));
        d.step(f, iflag, info, 1.0e-20, 2.220446049250313e-16);   // EPS, XTOL of Inversion.f90:79-80
        HIPCHK(hipMemcpyAsync(x, d.x, b, hipMemcpyDeviceToHost, d.st));
        HIPCHK(hipStreamSynchronize(d.st));
    });
}

int sfd_lbfgs_run(const sfd_lbfgs_opts *o, int n, float *x, const float *vpmask, const double *vpinit,
                  sfd_fctgdt fct, void *ctx, sfd_lbfgs_report *rep) {
    if (!o || n < 1 || !x || !fct) return SFD_EINVAL;
    return abi_call(nullptr, [&]() {
        LbfgsFct f = [&](const float *xx, float *g, double *fv) { return fct(ctx, xx, g, fv); };
        lbfgs_drive(*o, n, x, vpmask, vpinit, f, rep, nullptr);
    });
}

int sfd_lbfgsoc_step(const char *state_in, const char *state_out, const sfd_lbfgsoc_opts *o, int n, double *x,
                     double f, const double *g, int *iflag, int *info) {
    if (!o || !state_out || n < 1 || !x || !g || !iflag || !info || o->m < 1 || o->nparam < 1 ||
        o->nparam > n || !o->xmin || !o->xmax)
        return SFD_EINVAL;
    if (*iflag != 0 && *iflag != 1) return SFD_EINVAL;   // DIAGCO = .FALSE.: no IFLAG 2 re-entry
    if (*iflag == 1 && !state_in) return SFD_EINVAL;
    return abi_call(nullptr, [&]() {
        LbfgsDev d(n, o->m, 0.0, 0.0, nullptr, o->clip_type, 1, nullptr);
        d.nparam = o->nparam;
        d.pxmin.assign(o->xmin, o->xmin + o->nparam);
        d.pxmax.assign(o->xmax, o->xmax + o->nparam);
        d.stp1_opt = o->stp1_opt;
        d.step_init = o->step_init;
        d.xtrapf = 4.0;   // MCSRCHOC's extrapolation bound (nlinv_lbfgs.f:752)
        if (*iflag == 1) d.load_state(state_in);
        const size_t b = (size_t)n * sizeof(double);
        HIPCHK(hipMemcpyAsync(d.x, x, b, hipMemcpyHostToDevice, d.st));
        HIPCHK(hipMemcpyAsync(d.g, g, b, hipMemcpyHostToDevice, d.st));
        d.step(f, iflag, info, 1.0e-20, 2.220446049250313e-16);   // nlinv_read_array_params.f90:227-228
        HIPCHK(hipMemcpyAsync(x, d.x, b, hipMemcpyDeviceToHost, d.st));
        HIPCHK(hipStreamSynchronize(d.st));
        if (*iflag >= 0) d.save_state(state_out);
    });
}

int sfd_lbfgsoc_step_files(const char *dir, const sfd_lbfgsoc_opts *o, int n, double *x, double f,
                           const double *g, int *iflag, int *info) {
    if (!dir || !o || n < 1 || !x || !g || !iflag || !info || o->m < 1 || o->nparam < 1 || o->nparam > n ||
        !o->xmin || !o->xmax)
        return SFD_EINVAL;
    return abi_call(nullptr, [&]() {
        const std::string D = std::string(dir) + "/";
        int fl = lbx::read_flag(D + "Iflag.dat");
        int mi = lbx::read_flag(D + "Info.dat");
        // DIAGCO = .FALSE.: LBFGSOC never returns IFLAG 2, and resuming with it
        // would run the two-loop on a DIAG it did not compute
        if (fl == 2) throw EngineError(SFD_EINVAL, D + "Iflag.dat holds 2 (DIAGCO re-entry is not supported)");
        LbfgsDev d(n, o->m, 0.0, 0.0, nullptr, o->clip_type, 1, nullptr);
        d.nparam = o->nparam;
        d.pxmin.assign(o->xmin, o->xmin + o->nparam);
        d.pxmax.assign(o->xmax, o->xmax + o->nparam);
        d.stp1_opt = o->stp1_opt;
        d.step_init = o->step_init;
        d.xtrapf = 4.0;   // MCSRCHOC's extrapolation bound (nlinv_lbfgs.f:752)
        if (fl != 0) {
            // L_BFGS.f90 reads DIAG and W for any IFLAG but 0; LBFGSOC then reads
            // its scalars and, for IFLAG 1, resumes the line search (MCSRCHOC
            // reads its own file when INFO is not 0); other values start afresh
            lbx::read_vector(d, d.diag, D + "Diag.dat");
            lbx::read_work(d, D + "WorkArray.dat");
            lbx::read_lbfgs_dat(d, D + "LbfgsDat.dat", fl == 1);
            if (fl == 1 && d.info != 0) lbx::read_mcsrch_dat(d, D + "McsrchDat.dat");
        }
        d.on_lbfgs_write = [&]() { lbx::write_lbfgs_dat(d, D + "LbfgsDatOut.dat"); };
        d.on_mcsrch_write = [&]() { lbx::write_mcsrch_dat(d, D + "McsrchDatOut.dat"); };
        const size_t b = (size_t)n * sizeof(double);
        HIPCHK(hipMemcpyAsync(d.x, x, b, hipMemcpyHostToDevice, d.st));
        HIPCHK(hipMemcpyAsync(d.g, g, b, hipMemcpyHostToDevice, d.st));
        d.step(f, &fl, &mi, 1.0e-20, 2.220446049250313e-16);   // nlinv_read_array_params.f90:227-228
        HIPCHK(hipMemcpyAsync(x, d.x, b, hipMemcpyDeviceToHost, d.st));
        HIPCHK(hipStreamSynchronize(d.st));
        lbx::write_work(d, D + "WorkArrayOut.dat");
        lbx::write_vector(d, d.diag, D + "DiagOut.dat");
        ftn::ListWriter::save_text(D + "IflagOut.dat", ftn::i2_record(fl));
        ftn::ListWriter::save_text(D + "InfoOut.dat", ftn::i2_record(mi));
        *iflag = fl;
        *info = mi;
    });
}

int sfd_lbfgs_destroy(sfd_lbfgs_handle lb) {
    if (!lb) return SFD_EINVAL;
    delete lb;
    return SFD_OK;
}

}  // extern "C"
