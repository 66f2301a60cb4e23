// lbfgs_host.h — host control flow of the device L-BFGS (restates
// libs/libinv/lbfgs.f; see lbfgs_dev.h for the vector kernels).
//
// W(1:N) of the reference is the device vector `q`; the S and Y histories are
// device arrays of m vectors; the scalars W(N+1:N+M) (rho) and
// W(N+M+1:N+2M) (alpha) live on the host.  DIAG doubles as MCSRCH's WA
// exactly as in the reference call (lbfgs.f:486-488).
#pragma once
#include <functional>
#include <vector>

#include "common.h"
#include "lbfgs_api.h"

struct LbfgsDev {
    long n = 0;
    int m = 5;
    double xmin = 0, xmax = 0;
    int boundt = 0, type = 1;
    hipStream_t st = nullptr;
    double *x = nullptr, *g = nullptr, *q = nullptr, *S = nullptr, *Y = nullptr, *diag = nullptr, *xinit = nullptr;
    double *part = nullptr, *res = nullptr, *res_h = nullptr;
    float *mask = nullptr;
    std::vector<double> rho, alpha;
    // LBFGS SAVE state
    int iter = 0, nfun = 0, point = 0, maxfev = 20, info = 0, bound = 0, npt = 0, cp = 0, nfev = 0, finish = 0;
    double gnorm = 0, xnorm2 = 0, stp1 = 0, ftol = 1e-4, stp = 0, ys = 0, yy = 0;
    double gtol = 0.9, stpmin = 1e-20, stpmax = 1e20;   // block data LB2 (lbfgs.f:636-641)
    // extrapolation bound of the line search: 4.0D1 in MCSRCH (lbfgs.f:792),
    // 4.0D0 in the out-of-core MCSRCHOC (nlinv_lbfgs.f:752)
    double xtrapf = 40.0;
    // MCSRCH SAVE state
    int infoc = 0, brackt = 0, stage1 = 0;
    double xtol_ = 2.220446049250313e-16;
    double dginit = 0, dgtest = 0, finit = 0, width = 0, width1 = 0, stx = 0, fx = 0, dgx = 0, sty = 0, fy = 0,
           dgy = 0, stmin = 0, stmax = 0;
    // optional preconditioner for the two-loop's initial matrix (H0 q)
    void (*h0)(void *ctx, double *q, long n, hipStream_t st) = nullptr;
    void *h0_ctx = nullptr;
    bool h0_active = false;
    bool used_h0 = false;      // the current search direction came from the learned H0
    int h0_rejects = 0;        // learned directions replaced because they were not descent directions
    // out-of-core variant (LBFGSOC): per-parameter clip, its STP1 rule
    int nparam = 0;                      // > 0: FREEZE_XOC blocks instead of FREEZE_X
    std::vector<double> pxmin, pxmax;
    int stp1_opt = -1;                   // < 0: libinv's rule (lbfgs.f:333-338)
    double step_init = 999999.0;
    // Scratch scalars of LBFGSOC / MCSRCHOC that the reference's state files
    // carry although every use recomputes them first (nlinv_lbfgs.f:277-281,
    // :756-760): kept so that a state file read and written again keeps them.
    double sq = 0, yr = 0, beta = 0, xnorm = 0;
    int inmc = 0, iycn = 0, iscn = 0;
    double dg = 0, dgm = 0, dgxm = 0, dgym = 0, ftest1 = 0, fm = 0, fxm = 0, fym = 0;
    int mp = 1020, lp = 0;   // COMMON /LB5/ units (block data LB4, nlinv_lbfgs.f:981-986)
    // optional callbacks at the points where LBFGSOC / MCSRCHOC write their
    // state files (nlinv_lbfgs.f:509-516, :548-565; :854-861, :898-905)
    std::function<void()> on_lbfgs_write, on_mcsrch_write;
    // optional observer of every accepted (gradient, step) pair
    void (*on_pair)(void *ctx, const double *gk, const double *sk, const double *yk, long n, hipStream_t st) = nullptr;
    void *pair_ctx = nullptr;

    LbfgsDev(long n_, int m_, double xmin_, double xmax_, const float *mask_h, int boundt_, int type_,
             const double *xinit_h);
    ~LbfgsDev();
    void init(const float *mask_h, const double *xinit_h);
    void release();
    double dot(const double *a, const double *b);
    void axpby(double a, const double *xv, double b, double *y);
    void scale_copy(double a, const double *xv, double *y);
    void freeze(double *xv);
    double *Sv(int k) const { return S + (size_t)k * n; }
    double *Yv(int k) const { return Y + (size_t)k * n; }
    // One call of LBFGS(N,M,X,F,G,...) on the device vectors x, g.
    void step(double f, int *iflag, int *myinfo, double eps, double xtol);
    void mcsrch(double f, double *s);
    void two_loop(bool use_h0);
    int restart_plain();
    void mul_diag(double *v);   // v *= DIAG
    // the complete optimiser state (scalars, W = q/rho/alpha/S/Y, DIAG)
    void save_state(const char *path);
    void load_state(const char *path);
};

