// cnn.hip — the learned gradient preconditioner (no reference counterpart):
// a 4-layer 3x3 CNN (1 -> 16 -> 16 -> 16 -> 1) trained on the GPU from the
// (gradient, accepted step) pairs of the first L-BFGS iterations, then used
// as the initial inverse Hessian H0 of the L-BFGS two-loop recursion.  The
// misfit line search judges every preconditioned step; rejected steps fall
// back to plain L-BFGS from the same point.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "../../include/seisfd.h"
#include "cnn_api.h"
#include "cnn_kernels.h"
#include "common.h"
#include "lbfgs_host.h"

using namespace cnn;

namespace {

constexpr int kWgradBlocks = 256;   // 1024 wavefronts of partial weight gradients
constexpr int kLossBlocks = 256;

dim3 tiles_grid(long np) { return dim3((unsigned)((np + 63) / 64)); }   // 4 wavefronts x 16 pixels
dim3 stream_grid(long n) { return dim3((unsigned)std::min<long>((n + 255) / 256, 4096)); }

}  // namespace

// The network, its training state and its device buffers.
struct Cnn {
    int nz, nx, nimg_max;
    long P;
    hipStream_t st = nullptr;
    float *theta = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr, *real = nullptr;
    unsigned short *Wf = nullptr, *Wb = nullptr;
    float *bias = nullptr;
    unsigned short *a[4] = {nullptr, nullptr, nullptr, nullptr};   // a0 input, a1..a3 hidden
    unsigned short *dy = nullptr, *dA = nullptr, *dB = nullptr;
    float *y = nullptr, *x32 = nullptr, *t32 = nullptr, *part = nullptr;
    double *lpart = nullptr, *qmax = nullptr, *mpart = nullptr;
    int adam_t = 0;

    Cnn(int nz_, int nx_, int nimg, int seed) : nz(nz_), nx(nx_), nimg_max(nimg), P((long)nz_ * nx_) {
        if (nz < 1 || nx < 1 || nimg < 1) throw EngineError(SFD_EINVAL, "CNN needs nz, nx, nimg >= 1");
        HIPCHK(hipStreamCreate(&st));
        const long NP = P * nimg;
        for (float **p : {&theta, &grad, &m, &v, &real}) HIPCHK(hipMalloc(p, NPARAM * sizeof(float)));
        HIPCHK(hipMalloc(&Wf, LAYERS * WL * 2));
        HIPCHK(hipMalloc(&Wb, LAYERS * WL * 2));
        HIPCHK(hipMalloc(&bias, LAYERS * C * sizeof(float)));
        for (auto &p : a) HIPCHK(hipMalloc(&p, NP * C * 2));
        for (unsigned short **p : {&dy, &dA, &dB}) HIPCHK(hipMalloc(p, NP * C * 2));
        HIPCHK(hipMalloc(&y, NP * sizeof(float)));
        HIPCHK(hipMalloc(&x32, NP * sizeof(float)));
        HIPCHK(hipMalloc(&t32, NP * sizeof(float)));
        HIPCHK(hipMalloc(&part, (size_t)kWgradBlocks * 4 * NACC * C * C * sizeof(float)));
        HIPCHK(hipMalloc(&lpart, kLossBlocks * sizeof(double)));
        HIPCHK(hipMalloc(&qmax, 2 * sizeof(double)));
        HIPCHK(hipMalloc(&mpart, 256 * sizeof(double)));
        // He initialisation of the real weights; padding and biases zero
        std::vector<float> th(NPARAM, 0.f), re(NPARAM, 0.f);
        std::mt19937 rng((unsigned)seed);
        for (int l = 0; l < LAYERS; l++) {
            const int ci_n = l == 0 ? 1 : C, co_n = l == LAYERS - 1 ? 1 : C;
            std::normal_distribution<float> nd(0.f, std::sqrt((l == LAYERS - 1 ? 1.f : 2.f) / (TAPS * ci_n)));
            for (int t = 0; t < TAPS; t++)
                for (int co = 0; co < co_n; co++)
                    for (int ci = 0; ci < ci_n; ci++) {
                        const int k = l * PL + (t * C + co) * C + ci;
                        th[k] = nd(rng);
                        re[k] = 1.f;
                    }
            for (int co = 0; co < co_n; co++) re[l * PL + WL + co] = 1.f;
        }
        HIPCHK(hipMemcpy(real, re.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice));
        set_params(th.data());
    }
    ~Cnn() {
        for (void *p : {(void *)theta, (void *)grad, (void *)m, (void *)v, (void *)real, (void *)Wf, (void *)Wb,
                        (void *)bias, (void *)a[0], (void *)a[1], (void *)a[2], (void *)a[3], (void *)dy,
                        (void *)dA, (void *)dB, (void *)y, (void *)x32, (void *)t32, (void *)part, (void *)lpart,
                        (void *)qmax, (void *)mpart})
            if (p) (void)hipFree(p);
        if (st) (void)hipStreamDestroy(st);
    }
    void check_nimg(int nimg) const {
        if (nimg < 1 || nimg > nimg_max) throw EngineError(SFD_EINVAL, "image count outside the CNN's capacity");
    }
    void pack() {
        hipLaunchKernelGGL(pack_kernel, dim3((LAYERS * WL + 255) / 256), dim3(256), 0, st, (const float *)theta, Wf,
                           Wb, bias);
    }
    void set_params(const float *th_h) {
        std::vector<float> th(th_h, th_h + NPARAM), re(NPARAM);
        HIPCHK(hipMemcpy(re.data(), real, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
        for (int i = 0; i < NPARAM; i++) th[i] *= re[i];   // padding stays zero
        HIPCHK(hipMemcpy(theta, th.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(hipMemsetAsync(m, 0, NPARAM * sizeof(float), st));
        HIPCHK(hipMemsetAsync(v, 0, NPARAM * sizeof(float), st));
        adam_t = 0;
        pack();
        HIPCHK(hipStreamSynchronize(st));
    }
    void get_params(float *th_h) {
        HIPCHK(hipMemcpy(th_h, theta, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    }
    // a[0] holds the input; y receives the output (FP32, channel 0)
    void forward(int nimg) {
        const Geom G{nz, nx, nimg};
        const dim3 gr = tiles_grid(P * nimg), bl(256);
        for (int l = 0; l < LAYERS - 1; l++)
            hipLaunchKernelGGL((conv_kernel<true, false>), gr, bl, 0, st, G, (const unsigned short *)a[l],
                               (const unsigned short *)(Wf + l * WL), (const float *)(bias + l * C),
                               (const unsigned short *)nullptr, (void *)a[l + 1]);
        hipLaunchKernelGGL((conv_kernel<false, true>), gr, bl, 0, st, G, (const unsigned short *)a[LAYERS - 1],
                           (const unsigned short *)(Wf + (LAYERS - 1) * WL), (const float *)(bias + (LAYERS - 1) * C),
                           (const unsigned short *)nullptr, (void *)y);
    }
    // after forward(): MSE against t32, gradient of every parameter into grad
    double backward(int nimg) {
        const Geom G{nz, nx, nimg};
        const long NP = P * nimg;
        const dim3 gr = tiles_grid(NP), bl(256);
        hipLaunchKernelGGL(loss_kernel, dim3(kLossBlocks), bl, 0, st, NP, (const float *)y, (const float *)t32, dy,
                           lpart);
        const unsigned short *d = dy;
        unsigned short *nxt = dA;
        for (int l = LAYERS - 1; l >= 0; l--) {
            hipLaunchKernelGGL(wgrad_kernel, dim3(kWgradBlocks), bl, 0, st, G, (const unsigned short *)a[l], d, part);
            hipLaunchKernelGGL(wgrad_reduce_kernel, dim3((NACC * C * C + 255) / 256), bl, 0, st, (const float *)part,
                               kWgradBlocks * 4, grad + l * PL);
            if (l == 0) break;
            // input gradient of layer l, times the ReLU derivative of its input a[l]
            hipLaunchKernelGGL((conv_kernel<false, false>), gr, bl, 0, st, G, d,
                               (const unsigned short *)(Wb + l * WL), (const float *)nullptr,
                               (const unsigned short *)a[l], (void *)nxt);
            d = nxt;
            nxt = nxt == dA ? dB : dA;
        }
        std::vector<double> lp(kLossBlocks);
        HIPCHK(hipMemcpyAsync(lp.data(), lpart, kLossBlocks * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        double s = 0.0;
        for (double q : lp) s += q;
        return s / (double)NP;
    }
    void adam(float lr) {
        adam_t++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float bc1 = 1.f - std::pow(b1, (float)adam_t), bc2 = 1.f - std::pow(b2, (float)adam_t);
        hipLaunchKernelGGL(adam_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, st, NPARAM, theta,
                           (const float *)grad, m, v, (const float *)real, lr, b1, b2, 1e-8f, bc1, bc2);
        pack();
    }
    // x32 / t32 hold nimg input / target images
    void train(int nimg, int steps, float lr, double *first, double *last) {
        hipLaunchKernelGGL(input_f32_kernel, stream_grid(P * nimg), dim3(256), 0, st, P * nimg, (const float *)x32,
                           a[0]);
        for (int k = 0; k < steps; k++) {
            forward(nimg);
            const double l = backward(nimg);
            if (k == 0 && first) *first = l;
            adam(lr);
        }
        forward(nimg);
        const double l = backward(nimg);
        if (last) *last = l;
    }
    void maxabs(const double *q, long n, double *out) {
        hipLaunchKernelGGL(maxabs_partial_kernel, dim3(256), dim3(256), 0, st, n, q, mpart);
        hipLaunchKernelGGL(maxabs_final_kernel, dim3(1), dim3(256), 0, st, 256, (const double *)mpart, out);
    }
    // out = O(q) = max|q| (N(q/max|q|) - N(-q/max|q|)) / 2 * scale  (odd, degree-1 homogeneous)
    void apply_odd(const double *q, double *out, double scale) {
        maxabs(q, P, qmax);
        hipLaunchKernelGGL(input_pm_kernel, stream_grid(P), dim3(256), 0, st, P, q, (const double *)qmax, a[0]);
        forward(2);
        hipLaunchKernelGGL(h0_out_kernel, stream_grid(P), dim3(256), 0, st, P, (const float *)y,
                           (const double *)qmax, scale, out);
    }
};

// ---------------------------------------------------------------------------
// L-BFGS hooks: collect pairs, train once, serve H0, judge preconditioned steps.
struct CnnPrecond : LbfgsHooks {
    sfd_precond_opts o;
    const int nz, nx, r, nzc, nxc;   // model grid, coarsening factor, the network's grid
    const long Pf;                   // model cells
    Cnn net;
    LbfgsDev *d = nullptr;
    std::vector<double *> gk, sk, yk;
    double *tmp = nullptr, *tmp2 = nullptr, *tmp3 = nullptr, *gmax = nullptr;
    double *cq = nullptr, *cy = nullptr, *cin = nullptr, *ctg = nullptr;   // coarse-grid images (r > 1)
    float *illum = nullptr, *w = nullptr;
    bool trained = false, have_w = false;
    int pc_iterations = 0, fallbacks = 0;
    double train_seconds = 0, loss_first = 0, loss_last = 0;

    static int factor(const sfd_precond_opts &o) { return std::max(1, o.coarsen); }
    CnnPrecond(int nz_, int nx_, const sfd_precond_opts &opt)
        : o(opt), nz(nz_), nx(nx_), r(factor(opt)), nzc((nz_ + factor(opt) - 1) / factor(opt)),
          nxc((nx_ + factor(opt) - 1) / factor(opt)), Pf((long)nz_ * nx_),
          net(nzc, nxc, 2 * std::max(1, opt.train_after), opt.seed) {
        if (o.coarsen < 0 || r > 64) throw EngineError(SFD_EINVAL, "coarsen must be 0..64");
        HIPCHK(hipMalloc(&tmp, Pf * sizeof(double)));
        HIPCHK(hipMalloc(&tmp2, Pf * sizeof(double)));
        HIPCHK(hipMalloc(&tmp3, Pf * sizeof(double)));
        HIPCHK(hipMalloc(&illum, Pf * sizeof(float)));
        HIPCHK(hipMalloc(&w, Pf * sizeof(float)));
        HIPCHK(hipMalloc(&gmax, 2 * sizeof(double)));
        if (r > 1)
            for (double **p : {&cq, &cy, &cin, &ctg}) HIPCHK(hipMalloc(p, net.P * sizeof(double)));
    }
    ~CnnPrecond() override {
        for (double *p : gk) (void)hipFree(p);
        for (double *p : sk) (void)hipFree(p);
        for (double *p : yk) (void)hipFree(p);
        for (void *p : {(void *)tmp, (void *)tmp2, (void *)tmp3, (void *)gmax, (void *)illum, (void *)w, (void *)cq,
                        (void *)cy, (void *)cin, (void *)ctg})
            if (p) (void)hipFree(p);
    }
    static void pair_thunk(void *ctx, const double *g, const double *s, const double *y, long n, hipStream_t st) {
        CnnPrecond *c = (CnnPrecond *)ctx;
        if (c->trained || (int)c->gk.size() >= c->o.train_after) return;
        double *gc = nullptr, *sc = nullptr, *yc = nullptr;
        HIPCHK(hipMalloc(&gc, n * sizeof(double)));
        HIPCHK(hipMalloc(&sc, n * sizeof(double)));
        HIPCHK(hipMalloc(&yc, n * sizeof(double)));
        HIPCHK(hipMemcpyAsync(gc, g, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpyAsync(sc, s, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpyAsync(yc, y, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
        c->gk.push_back(gc);
        c->sk.push_back(sc);
        c->yk.push_back(yc);
    }
    static void h0_thunk(void *ctx, double *q, long n, hipStream_t st) { ((CnnPrecond *)ctx)->apply(q); }

    // the coarse-grid smoother P R alone is usable before the network is trained
    bool smoother_only() const { return r > 1 && !o.use_net; }
    void attach(LbfgsDev &dd) override {
        if (dd.n != Pf) throw EngineError(SFD_EINVAL, "preconditioner image size differs from the model size");
        d = &dd;
        dd.h0 = &h0_thunk;
        dd.h0_ctx = this;
        dd.on_pair = &pair_thunk;
        dd.pair_ctx = this;
        dd.h0_active = o.illum_h0 != 0 || smoother_only();   // a network-free operator is usable from the start
    }
    bool wants_illumination() const override { return o.illum_h0 != 0; }
    // W = (illum + 1e-4 max illum)^(-1/2): the pseudo-Hessian diagonal of the
    // summed source illumination, stabilised as compute_fct_gdt does (:338)
    void set_illumination(const float *il, long n) override {
        if (n != Pf) return;
        float mx = 0.f;
        for (long i = 0; i < n; i++) mx = std::max(mx, il[i]);
        HIPCHK(hipMemcpyAsync(illum, il, n * sizeof(float), hipMemcpyHostToDevice, net.st));
        hipLaunchKernelGGL(illum_weight_kernel, stream_grid(n), dim3(256), 0, net.st, n, (const float *)illum,
                           1e-4f * mx, w);
        HIPCHK(hipStreamSynchronize(net.st));
        have_w = true;
    }
    bool net_on() const { return o.use_net && trained; }
    void restrict_to(const double *f, double *c) {
        hipLaunchKernelGGL(restrict_kernel, stream_grid(net.P), dim3(256), 0, net.st, f, nz, nx, r, nzc, nxc, c);
    }
    // M(src) = W P N(R W src) * scale, with W = 1 without illumination, N =
    // identity without the network and P = R = identity when r = 1
    void op(const double *src, double *dst, double scale) {
        const float *wp = (o.illum_h0 && have_w) ? w : nullptr;
        hipLaunchKernelGGL(scale_w_kernel, stream_grid(Pf), dim3(256), 0, net.st, Pf, wp, src, 1.0, tmp3);
        const double *mid = tmp3;
        if (r > 1 && (net_on() || smoother_only())) {
            restrict_to(tmp3, cq);
            if (net_on()) net.apply_odd(cq, cy, 1.0);
            hipLaunchKernelGGL(prolong_kernel, stream_grid(Pf), dim3(256), 0, net.st,
                               (const double *)(net_on() ? cy : cq), nzc, nxc, r, nz, nx, dst);
            mid = dst;
        } else if (net_on()) {
            net.apply_odd(tmp3, dst, 1.0);
            mid = dst;
        }
        hipLaunchKernelGGL(scale_w_kernel, stream_grid(Pf), dim3(256), 0, net.st, Pf, wp, mid, scale, dst);
        HIPCHK(hipStreamSynchronize(net.st));
    }
    void train() {
        HIPCHK(hipStreamSynchronize(d->st));
        const auto t0 = std::chrono::steady_clock::now();
        const int K = (int)gk.size();
        if (o.target == 1) {
            // s_k <- sum_{j >= k} s_j: the accepted update from x_k to the newest model
            for (int k = K - 2; k >= 0; k--)
                hipLaunchKernelGGL(accumulate_kernel, stream_grid(Pf), dim3(256), 0, net.st, Pf,
                                   (const double *)sk[k + 1], sk[k]);
        }
        // target 2 (secant pairs): N maps y_k to s_k, in the coordinates of the
        // operator W N(W .): inputs W y_k, targets s_k / W
        const bool secant = o.target == 2;
        const float *wp = (o.illum_h0 && have_w) ? w : nullptr;
        const long P = net.P;
        for (int k = 0; k < K; k++) {
            const double *in = gk[k], *tg = sk[k];
            if (secant) {
                hipLaunchKernelGGL(scale_w_kernel, stream_grid(Pf), dim3(256), 0, net.st, Pf, wp,
                                   (const double *)yk[k], 1.0, tmp);
                hipLaunchKernelGGL(unscale_w_kernel, stream_grid(Pf), dim3(256), 0, net.st, Pf, wp,
                                   (const double *)sk[k], tmp2);
                in = tmp;
                tg = tmp2;
            }
            if (r > 1) {   // the network learns on the coarse grid
                restrict_to(in, cin);
                restrict_to(tg, ctg);
                in = cin;
                tg = ctg;
            }
            net.maxabs(in, P, gmax);
            net.maxabs(tg, P, gmax + 1);
            hipLaunchKernelGGL(pair_kernel, stream_grid(P), dim3(256), 0, net.st, P, in, (const double *)gmax, tg,
                               (const double *)(gmax + 1), net.x32 + k * P, net.x32 + (K + k) * P, net.t32 + k * P,
                               net.t32 + (K + k) * P, secant ? 1.0 : -1.0);
        }
        net.train(2 * K, o.train_steps, o.lr, &loss_first, &loss_last);
        HIPCHK(hipStreamSynchronize(net.st));
        train_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        trained = true;
    }
    void accepted(LbfgsDev &dd, int) override {
        if (dd.used_h0) pc_iterations++;
        if (o.use_net && !trained && (int)gk.size() >= o.train_after) {
            train();
            dd.h0_active = fallbacks < o.max_fallbacks;
        }
    }
    int after_step(LbfgsDev &dd, int iflag) override {
        if (!dd.used_h0) return iflag;
        const bool failed = iflag < 0 || (iflag == 1 && dd.info == -1 && dd.nfev >= o.max_evals);
        if (!failed) return iflag;
        if (++fallbacks >= o.max_fallbacks) dd.h0_active = false;
        return dd.restart_plain();
    }
    int rejected_trial(LbfgsDev &dd) override {
        if (++fallbacks >= o.max_fallbacks) dd.h0_active = false;
        return dd.restart_plain();
    }
    // H0 q = gamma O(q), gamma = ys / (y . O(y)) from the newest pair (the
    // scaling of the plain H0 = ys/yy I along y); plain H0 if y . O(y) <= 0.
    void apply(double *q) {
        HIPCHK(hipStreamSynchronize(d->st));
        const double *yv = d->Yv(d->npt);
        op(yv, tmp, 1.0);
        const double yoy = d->dot(yv, tmp);
        if (!(yoy > 0.0) || !std::isfinite(yoy)) {
            d->used_h0 = false;
            d->mul_diag(q);   // DIAG = ys/yy: the plain H0
            return;
        }
        op(q, tmp2, d->ys / yoy);
        if (!std::isfinite(d->dot(tmp2, tmp2))) {
            d->used_h0 = false;
            d->mul_diag(q);
            return;
        }
        HIPCHK(hipMemcpyAsync(q, tmp2, Pf * sizeof(double), hipMemcpyDeviceToDevice, net.st));
        HIPCHK(hipStreamSynchronize(net.st));
    }
    void report(sfd_lbfgs_report *rep) override {
        rep->pc_iterations = pc_iterations;
        rep->pc_fallbacks = fallbacks;
        rep->pc_rejects = d ? d->h0_rejects : 0;
        rep->pc_train_seconds = train_seconds;
        rep->pc_loss_first = loss_first;
        rep->pc_loss_last = loss_last;
    }
};

LbfgsHooks *make_cnn_precond(int nz, int nx, const sfd_precond_opts &o) { return new CnnPrecond(nz, nx, o); }

// ---------------------------------------------------------------------------
struct sfd_cnn {
    Cnn *net = nullptr;
    ~sfd_cnn() { delete net; }
};

extern "C" {

int sfd_cnn_create(sfd_cnn_handle *c, int nz, int nx, int nimg_max, int seed) {
    if (!c) return SFD_EINVAL;
    *c = nullptr;
    return abi_call(nullptr, [&]() {
        std::unique_ptr<sfd_cnn> h(new sfd_cnn());
        h->net = new Cnn(nz, nx, nimg_max, seed);
        *c = h.release();
    });
}

int sfd_cnn_destroy(sfd_cnn_handle c) {
    if (!c) return SFD_EINVAL;
    delete c;
    return SFD_OK;
}

int sfd_cnn_get_params(sfd_cnn_handle c, float *theta) {
    if (!c || !theta) return SFD_EINVAL;
    return abi_call(nullptr, [&]() {
        c->net->get_params(theta);
    });
}

int sfd_cnn_set_params(sfd_cnn_handle c, const float *theta) {
    if (!c || !theta) return SFD_EINVAL;
    return abi_call(nullptr, [&]() {
        c->net->set_params(theta);
    });
}

int sfd_cnn_forward(sfd_cnn_handle c, int nimg, const float *x, float *y) {
    if (!c || !x || !y) return SFD_EINVAL;
    return abi_call(nullptr, [&]() {
        Cnn &n = *c->net;
        n.check_nimg(nimg);
        const long NP = n.P * nimg;
        HIPCHK(hipMemcpyAsync(n.x32, x, NP * sizeof(float), hipMemcpyHostToDevice, n.st));
        hipLaunchKernelGGL(input_f32_kernel, stream_grid(NP), dim3(256), 0, n.st, NP, (const float *)n.x32, n.a[0]);
        n.forward(nimg);
        HIPCHK(hipMemcpyAsync(y, n.y, NP * sizeof(float), hipMemcpyDeviceToHost, n.st));
        HIPCHK(hipStreamSynchronize(n.st));
    });
}

int sfd_cnn_loss_grad(sfd_cnn_handle c, int nimg, const float *x, const float *target, double *loss, float *grad) {
    if (!c || !x || !target) return SFD_EINVAL;
    return abi_call(nullptr, [&]() {
        Cnn &n = *c->net;
        n.check_nimg(nimg);
        const long NP = n.P * nimg;
        HIPCHK(hipMemcpyAsync(n.x32, x, NP * sizeof(float), hipMemcpyHostToDevice, n.st));
        HIPCHK(hipMemcpyAsync(n.t32, target, NP * sizeof(float), hipMemcpyHostToDevice, n.st));
        hipLaunchKernelGGL(input_f32_kernel, stream_grid(NP), dim3(256), 0, n.st, NP, (const float *)n.x32, n.a[0]);
        n.forward(nimg);
        const double l = n.backward(nimg);
        if (loss) *loss = l;
        if (grad) HIPCHK(hipMemcpy(grad, n.grad, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int sfd_cnn_train(sfd_cnn_handle c, int nimg, const float *x, const float *target, int steps, float lr,
                  double *loss_first, double *loss_last) {
    if (!c || !x || !target || steps < 0) return SFD_EINVAL;
    return abi_call(nullptr, [&]() {
        Cnn &n = *c->net;
        n.check_nimg(nimg);
        const long NP = n.P * nimg;
        HIPCHK(hipMemcpyAsync(n.x32, x, NP * sizeof(float), hipMemcpyHostToDevice, n.st));
        HIPCHK(hipMemcpyAsync(n.t32, target, NP * sizeof(float), hipMemcpyHostToDevice, n.st));
        n.train(nimg, steps, lr, loss_first, loss_last);
        HIPCHK(hipStreamSynchronize(n.st));
    });
}

}  // extern "C"
