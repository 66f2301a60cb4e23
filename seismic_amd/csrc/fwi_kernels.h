// fwi_kernels.h — device kernels of the FWI misfit / gradient assembly
// (progs/afwi/compute_function_gradient.f90:208-350).
//
// FP64 is used for the misfit and the illumination norm, as the reference
// does (Norm.f90:28 dprod; compute_function_gradient.f90:338).  Reductions
// are fixed-order trees, so results are run-to-run deterministic.
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

// Compute_OF_RES_ADJ (Objective_Function_Residual_Adjsource.f90:32-38), L2:
//   r = w (dobs - dmod); f_trace = sum(dprod(r, r)) / 2; resid = r; adj = w r.
// The adjoint source is written straight into the shot's injection-trace
// buffer used by the reverse-time pass.
__global__ void residual_kernel(const TraceRef *__restrict__ refs, const ShotDev *__restrict__ shots,
                                const float *__restrict__ dobs, const float *__restrict__ mute,
                                float *__restrict__ resid, float *const *__restrict__ adj_base,
                                double *__restrict__ ftrace) {
    __shared__ double sh[256];
    const TraceRef R = refs[blockIdx.x];
    const ShotDev &S = shots[R.shot];
    const float *dmod = S.traces + (size_t)R.rec * R.nt;
    float *adj = adj_base[R.shot] + (size_t)R.rec * R.nt;
    double acc = 0.0;
    for (int it = threadIdx.x; it < R.nt; it += blockDim.x) {
        const long q = R.off + it;
        const float w = mute ? mute[q] : 1.0f;
        const float r = w * (dobs[q] - dmod[it]);
        resid[q] = r;
        adj[it] = w * r;
        acc += (double)r * (double)r;
    }
    const double s = block_sum_d(acc, sh);
    if (threadIdx.x == 0) ftrace[blockIdx.x] = s / 2;
}

// illu = illu**illupow; grad = 2 grad / scaling (compute_function_gradient.f90:332-337);
// per-block max(illu) and sum(dprod(illu, illu)) partials.
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
        grad[q] = 2.f * grad[q] / scal;
        m = q == blockIdx.x * (size_t)blockDim.x + threadIdx.x ? il : fmaxf(m, il);
        acc += (double)il * (double)il;
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
