// lbfgs_dev.h — L-BFGS of libs/libinv/lbfgs.f (LBFGS :49-558, MCSRCH
// :652-982, MCSTEP :985-1210, FREEZE_X :1-47) with every O(N) vector
// operation on the device in FP64 and the scalar control flow (the
// reference's SAVE variables and GOTO re-entry points) on the host.
//
// The reference's OpenMP REDUCTION dot products become fixed-order
// two-pass device reductions (deterministic run to run).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

namespace sfd {

constexpr int kDotBlocks = 512;

__global__ void dot_partial_kernel(const double *__restrict__ a, const double *__restrict__ b, long n,
                                   double *__restrict__ part) {
    __shared__ double sh[256];
    const long chunk = (n + gridDim.x - 1) / gridDim.x;
    const long lo = blockIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
    double acc = 0.0;
    for (long i = lo + threadIdx.x; i < hi; i += blockDim.x) acc += a[i] * b[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void dot_final_kernel(const double *__restrict__ part, int np, double *__restrict__ out) {
    __shared__ double sh[kDotBlocks];
    for (int i = threadIdx.x; i < kDotBlocks; i += blockDim.x) sh[i] = i < np ? part[i] : 0.0;
    __syncthreads();
    for (int s = kDotBlocks / 2; s > 0; s >>= 1) {
        for (int i = threadIdx.x; i < s; i += blockDim.x) sh[i] += sh[i + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// y = a*x + b*y  (general two-vector update used by every LBFGS loop)
__global__ void axpby_kernel(long n, double a, const double *__restrict__ x, double b, double *__restrict__ y) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = a * x[i] + b * y[i];
}
// y = x  (b == 0 above would still read y; this one does not)
__global__ void copy_scaled_kernel(long n, double a, const double *__restrict__ x, double *__restrict__ y) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = a * x[i];
}
// y = x * d (elementwise)
__global__ void mul_kernel(long n, const double *__restrict__ d, double *__restrict__ y) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = d[i] * y[i];
}
__global__ void fill_kernel(long n, double v, double *__restrict__ y) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] = v;
}
// y = g - w  (lbfgs.f:507)
__global__ void sub_kernel(long n, const double *__restrict__ g, const double *__restrict__ w, double *__restrict__ y) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = g[i] - w[i];
}
// FREEZE_X (lbfgs.f:1-47)
__global__ void freeze_kernel(long n, double xmin, double xmax, const float *__restrict__ mask, int type,
                              double *__restrict__ x, const double *__restrict__ xinit) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double v = x[i];
        const float mk = mask ? mask[i] : 1.0f;
        if (type || (mk != 0.f && mk != 2.f)) {
            v = fmax(xmin, v);
            v = fmin(xmax, v);
        }
        if (!type && (mk == 0.f || mk == 2.f)) v = xinit[i];
        x[i] = v;
    }
}

// FREEZE_XOC (nlinv_lbfgs.f:1-22) on one parameter block x[off, off + len)
__global__ void clip_block_kernel(long off, long len, double lo, double hi, double *__restrict__ x) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < len; i += (long)gridDim.x * blockDim.x) {
        double v = x[off + i];
        v = fmax(lo, v);
        v = fmin(hi, v);
        x[off + i] = v;
    }
}

}  // namespace sfd
