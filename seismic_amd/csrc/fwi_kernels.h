// fwi_kernels.h — device kernels of the FWI misfit / gradient assembly
// (progs/afwi/compute_function_gradient.f90:208-350).
//
// FP64 is used for the L2 misfit and the illumination norm, as the reference
// does (Norm.f90:28 dprod; compute_function_gradient.f90:338).  Each trace's
// misfit is summed in time order, as the reference's SUM does, so it is the
// reference's value bit for bit; other reductions are fixed-order trees, so
// results are run-to-run deterministic.
#pragma once
#include <hip/hip_runtime.h>

namespace sfd {

// Per-trace table of a batch: which shot, which receiver, offset of the trace
// in the concatenated (nt, ntr) blocks.
struct TraceRef {
    int shot, rec, nt;
    long off;
};

__device__ __forceinline__ double block_sum_d(double v, double *sh) {
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (tid < s) sh[tid] = sh[tid] + sh[tid + s];
        __syncthreads();
    }
    return sh[0];
}

// Compute_OF_RES_ADJ (Objective_Function_Residual_Adjsource.f90:32-38) in
// four steps, one block (or thread) per trace of the batch:
//   residual_kernel   r = w (dobs - dmod) -> resid; w r -> the shot's
//                     injection-trace buffer (the adjoint source before the
//                     norm); max|r| per trace (FP32, order-free).
//   norm_thresh_kernel  the THRESH each trace sees: one per shot, passed from
//                     trace to trace and set from the first trace whose
//                     max|r| is not 0 when data_threshold is 0 (Norm.f90:36-40).
//   misfit_kernel     fct_compute per trace, one thread per trace summing in
//                     time order: FP64 products for L2 (dprod), FP32 terms
//                     in FP32 otherwise (SUM of a default-real array).
//   adjsrc_kernel     gdt_compute on the adjoint source (not for L2).
__global__ void residual_kernel(const TraceRef *__restrict__ refs, const ShotDev *__restrict__ shots,
                                const float *__restrict__ dobs, const float *__restrict__ mute,
                                float *__restrict__ resid, float *const *__restrict__ adj_base,
                                float *__restrict__ tmax) {
    __shared__ float sh[256];
    const TraceRef R = refs[blockIdx.x];
    const ShotDev &S = shots[R.shot];
    const float *dmod = S.traces + (size_t)R.rec * R.nt;
    float *adj = adj_base[R.shot] + (size_t)R.rec * R.nt;
    float m = 0.f;
    for (int it = threadIdx.x; it < R.nt; it += blockDim.x) {
        const long q = R.off + it;
        const float w = mute ? mute[q] : 1.0f;
        const float r = w * (dobs[q] - dmod[it]);
        resid[q] = r;
        adj[it] = w * r;
        m = fmaxf(m, fabsf(r));
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0) tmax[blockIdx.x] = sh[0];
}

__device__ __forceinline__ bool norm_uses_thresh(int norm) {
    return norm == SFD_NORM_HUBER || norm == SFD_NORM_CAUCHY || norm == SFD_NORM_HYPERBOLIC;
}

// one thread per shot of the wave; shot s owns traces [tbeg[s], tbeg[s+1])
__global__ void norm_thresh_kernel(int nshots, const int *__restrict__ tbeg, const float *__restrict__ tmax,
                                   int norm, float thresh0, float *__restrict__ thr) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshots) return;
    float t = thresh0;
    for (int k = tbeg[s]; k < tbeg[s + 1]; k++) {
        if (norm_uses_thresh(norm) && t == 0.f) t = tmax[k];
        thr[k] = t;
    }
}

__global__ void misfit_kernel(int ntr, const TraceRef *__restrict__ refs, const float *__restrict__ resid,
                              const float *__restrict__ thr, int norm, double *__restrict__ ftrace) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ntr) return;
    const TraceRef R = refs[k];
    const float *r = resid + R.off;
    if (norm == SFD_NORM_L2 || norm == 0) {
        double f = 0.0;
#pragma unroll 8
        for (int it = 0; it < R.nt; it++) {
            const double d = (double)r[it];
            f = f + d * d;
        }
        ftrace[k] = f / 2;
        return;
    }
    const float t = thr[k];
    float acc = 0.f;
#pragma unroll 8
    for (int it = 0; it < R.nt; it++) {
        const float x = r[it], a = fabsf(x);
        float term;
        if (norm == SFD_NORM_L1) {
            term = a;
        } else if (norm == SFD_NORM_HUBER) {
            term = a > t ? a - t / 2.0f : x * x / (2.0f * t);
        } else if (norm == SFD_NORM_CAUCHY) {
            const float q = x / t;
            term = logf(1.0f + q * q) / 2.0f;
        } else {
            term = sqrtf(t * t + x * x) - t;
        }
        acc = acc + term;
    }
    ftrace[k] = (double)acc;
}

__global__ void adjsrc_kernel(const TraceRef *__restrict__ refs, float *const *__restrict__ adj_base,
                              const float *__restrict__ thr, int norm) {
    const TraceRef R = refs[blockIdx.x];
    float *adj = adj_base[R.shot] + (size_t)R.rec * R.nt;
    const float t = thr[blockIdx.x];
    for (int it = threadIdx.x; it < R.nt; it += blockDim.x) {
        const float x = adj[it];
        float y;
        if (norm == SFD_NORM_L1) {
            y = x < 0.f ? -1.f : (x > 0.f ? 1.f : 0.f);
        } else if (norm == SFD_NORM_HUBER) {   // max(-1, min(1, x/t)); NaN stays NaN
            const float q = x / t;
            y = q > 1.f ? 1.f : (q < -1.f ? -1.f : q);
        } else if (norm == SFD_NORM_CAUCHY) {
            y = x / (t * t + x * x);
        } else {
            y = (x / t) / sqrtf(1.0f + x * x / (t * t));
        }
        adj[it] = y;
    }
}

// Mute mask of one trace per block (Mute.f90:171-219): top(i) is 0 before
// the ramp, ramp_top[i - itop + nramp] on [itop - nramp, itop] and 1 after
// itop; bottom(i) is 1 before ibot + 1, ramp_bot[i - ibot - 1] on
// [ibot + 1, ibot + nramp] and 0 from ibot + 1 + nramp on (1-based samples).
// The ramps (sin^2 / cos^2 of k pi / (2 nramp)) and the onsets come from the
// host, which evaluates them with the reference's FP32 expressions.
__global__ void mute_kernel(const TraceRef *__restrict__ refs, const int2 *__restrict__ onset, int nramp,
                            const float *__restrict__ ramp_top, const float *__restrict__ ramp_bot,
                            float *__restrict__ mask) {
    const TraceRef R = refs[blockIdx.x];
    const int itop = onset[blockIdx.x].x, ibot = onset[blockIdx.x].y;
    for (int i0 = threadIdx.x; i0 < R.nt; i0 += blockDim.x) {
        const int i = i0 + 1;
        const float top = i > itop ? 1.f : (i >= itop - nramp ? ramp_top[i - itop + nramp] : 0.f);
        const float bot = i >= ibot + 1 + nramp ? 0.f : (i >= ibot + 1 ? ramp_bot[i - ibot - 1] : 1.f);
        mask[R.off + i0] = top * bot;
    }
}

// illu = illu**illupow; grad = 2 grad / scaling (compute_function_gradient.f90:332-337);
// per-block max(illu) and sum(dprod(illu, illu)) partials.  A gradient value
// that is not finite makes its block's sum NaN, so the caller's finiteness
// check on the illumination norm reports it.
__global__ void fwi_scale_kernel(float *__restrict__ grad, float *__restrict__ illu, size_t n, float illupow,
                                 float scal, float *__restrict__ bmax, double *__restrict__ bsum) {
    __shared__ double sh[256];
    __shared__ float shm[256];
    float m = 0.f;
    double acc = 0.0;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        float il = illu[q];
        if (illupow != 1.0f) il = powf(il, illupow);
        illu[q] = il;
        const float gq = 2.f * grad[q] / scal;
        grad[q] = gq;
        m = q == blockIdx.x * (size_t)blockDim.x + threadIdx.x ? il : fmaxf(m, il);
        acc += (double)il * (double)il;
        if (!isfinite(gq)) acc = __builtin_nan("");
    }
    const double s = block_sum_d(acc, sh);
    shm[threadIdx.x] = m;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) shm[threadIdx.x] = fmaxf(shm[threadIdx.x], shm[threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        bmax[blockIdx.x] = shm[0];
        bsum[blockIdx.x] = s;
    }
}

// illu = (illu + max/10000) / rn (in double, stored FP32); grad = grad/illu; grad *= mask
// (compute_function_gradient.f90:338-341)
__global__ void fwi_normalize_kernel(float *__restrict__ grad, float *__restrict__ illu, size_t n, float mx,
                                     double rn, const float *__restrict__ mask) {
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const float il = (float)((double)(illu[q] + mx / 10000.f) / rn);
        illu[q] = il;
        float gq = grad[q] / il;
        if (mask) gq = gq * mask[q];
        grad[q] = gq;
    }
}

__global__ void mask_kernel(float *__restrict__ grad, size_t n, const float *__restrict__ mask) {
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        grad[q] = grad[q] * mask[q];
}

// Smoothing.f90:140-227 triangle (boxconv twice + end folds), one line per
// thread.  Line l starts at in + l*lstride, elements `stride` apart.  pp/qq
// are per-line scratch (n+nr, n+2nr floats), laid out interleaved across
// lines (element k of line l at k*nlines + l) so neighbouring threads touch
// neighbouring addresses.
__global__ void triangle_kernel(const float *__restrict__ in, float *__restrict__ out, int nlines, int n,
                                size_t lstride, size_t stride, int nr, float *__restrict__ pp,
                                float *__restrict__ qq) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nlines) return;
    const float *u = in + (size_t)l * lstride;
    float *v = out + (size_t)l * lstride;
#define PP(k) pp[(size_t)(k) * nlines + l]
#define QQ(k) qq[(size_t)(k) * nlines + l]
    if (n == 1) {
        v[0] = u[0];
        return;
    }
    // boxconv(nr, n, qq=u, pp): running sum b, y(i) = b(i) - b(i-nr), /nr
    {
        const int ny = n + nr - 1;
        float b = 0.f;
        // bb(i) for i <= n: prefix sums; i > n: constant.  yy(i) = bb(i) (i<=nr) else bb(i)-bb(i-nr)
        for (int i = 1; i <= ny; i++) {
            b = i == 1 ? u[0] : (i <= n ? b + u[(size_t)(i - 1) * stride] : b);
            PP(i - 1) = b;   // store bb first
        }
        for (int i = ny; i >= 1; i--) {
            const float y = i <= nr ? PP(i - 1) : PP(i - 1) - PP(i - nr - 1);
            QQ(i - 1) = y / (float)nr;   // stash y in qq temporarily
        }
        for (int i = 1; i <= ny; i++) PP(i - 1) = QQ(i - 1);
    }
    // boxconv(nr, np, pp, qq)
    {
        const int np = nr + n - 1, ny = np + nr - 1;
        float b = 0.f;
        for (int i = 1; i <= ny; i++) {
            b = i == 1 ? PP(0) : (i <= np ? b + PP(i - 1) : b);
            QQ(i - 1) = b;
        }
        for (int i = ny; i >= 1; i--) {
            const float y = i <= nr ? QQ(i - 1) : QQ(i - 1) - QQ(i - nr - 1);
            QQ(i - 1) = y / (float)nr;
        }
    }
    for (int i = 1; i <= n; i++) {
        float t = QQ(i + nr - 2);
        if (i <= nr - 1) t = t + QQ(nr - i - 1);   // fold back the near end
        v[(size_t)(i - 1) * stride] = t;
    }
    for (int i = 1; i <= nr - 1; i++) {   // fold back the far end
        const size_t o = (size_t)(n - i) * stride;
        v[o] = v[o] + QQ(n + (nr - 1) + i - 1);
    }
#undef PP
#undef QQ
}

}  // namespace sfd
