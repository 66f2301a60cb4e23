// cnn_kernels.h — the learned gradient preconditioner's network: four 3x3
// convolutions (1 -> 16 -> 16 -> 16 -> 1 channels, ReLU after the first
// three), BF16 weights and activations, FP32 accumulation on the matrix cores
// (v_mfma_f32_16x16x16_bf16), and its training kernels (input and weight
// gradients, MSE loss, Adam).
//
// Tensors hold 16 channels per pixel ([pixel][16], BF16); a 1-channel tensor
// is channel 0 with the other 15 channels zero, and the first and last
// layers are 16x16 layers whose weights outside the real 1-channel slice are
// held at zero.  Pixels are numbered as the model is stored: p = iz + nz*ix
// (column-major), with a stack of images following one another.
#pragma once
#include <hip/hip_runtime.h>

namespace cnn {

typedef short short4v __attribute__((ext_vector_type(4)));
typedef float float4m __attribute__((ext_vector_type(4)));

constexpr int C = 16;                 // channels per pixel
constexpr int TAPS = 9;               // 3x3 stencil: tap t -> (dz, dx) = (t%3 - 1, t/3 - 1)
constexpr int LAYERS = 4;
constexpr int WL = TAPS * C * C;      // weights of one layer, [t][co][ci]
constexpr int PL = WL + C;            // + biases
constexpr int NPARAM = LAYERS * PL;
constexpr int NACC = TAPS + 1;        // weight-gradient accumulators: 9 taps + bias

// FP32 -> BF16, round to nearest even (inputs are finite)
__device__ __forceinline__ unsigned short f2bf(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

struct Geom {
    int nz, nx, nimg;
};

__device__ __forceinline__ long npix(const Geom &G) { return (long)G.nz * G.nx * G.nimg; }

// Neighbour of pixel p at tap t, or -1 outside its image (zero padding).
__device__ __forceinline__ long nbr(const Geom &G, long p, int t) {
    const long P = (long)G.nz * G.nx;
    const long img = p / P, q = p - img * P;
    const int ix = (int)(q / G.nz), iz = (int)(q - (long)ix * G.nz);
    const int jz = iz + t % 3 - 1, jx = ix + t / 3 - 1;
    if (jz < 0 || jz >= G.nz || jx < 0 || jx >= G.nx) return -1;
    return img * P + jz + (long)G.nz * jx;
}

// out[p][co] = act(bias[co] + sum_t sum_ci in[nbr(p,t)][ci] * W[t][co][ci]),
// optionally multiplied by (mask[p][co] > 0) (ReLU derivative in the
// backward pass).  One wavefront per 16-pixel tile, one MFMA per tap:
//   A (16 pixels x 16 in-channels): lane l holds pixel l%16, channels 4(l/16)..+3
//   B (16 in x 16 out-channels):     lane l holds out-channel l%16, in-channels 4(l/16)..+3
//   D (16 pixels x 16 out-channels): lane l holds pixels 4(l/16)..+3, out-channel l%16
// OUT_F32 writes channel 0 only, as FP32 [pixel].
template <bool RELU, bool OUT_F32>
__global__ void __launch_bounds__(256) conv_kernel(Geom G, const unsigned short *__restrict__ in,
                                                   const unsigned short *__restrict__ W,
                                                   const float *__restrict__ bias,
                                                   const unsigned short *__restrict__ mask, void *__restrict__ out) {
    const long NP = npix(G);
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const long p0 = ((long)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    if (p0 >= NP) return;   // whole wavefront
    const long pa = p0 + r;
    float4m acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < TAPS; t++) {
        short4v a = {0, 0, 0, 0};
        if (pa < NP) {
            const long q = nbr(G, pa, t);
            if (q >= 0) a = *(const short4v *)(in + q * C + 4 * g);
        }
        const short4v b = *(const short4v *)(W + (t * C + r) * C + 4 * g);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    const float bs = bias ? bias[r] : 0.f;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const long p = p0 + 4 * g + j;
        if (p >= NP) continue;
        float v = acc[j] + bs;
        if (RELU) v = fmaxf(v, 0.f);
        if (mask && !(bf2f(mask[p * C + r]) > 0.f)) v = 0.f;
        if (OUT_F32) {
            if (r == 0) ((float *)out)[p] = v;
        } else {
            ((unsigned short *)out)[p * C + r] = f2bf(v);
        }
    }
}

// Partial weight gradients of one layer: for every wavefront w,
//   part[w][t][co][ci] = sum over its tiles of sum_p a[nbr(p,t)][ci] * dout[p][co]
// and accumulator t = 9 holds sum_p dout[p][co] (bias gradient) in ci = 0.
//   A (16 out-channels x 16 pixels): lane l holds channel l%16, pixels 4(l/16)..+3
//   B (16 pixels x 16 in-channels):  lane l holds in-channel l%16, pixels 4(l/16)..+3
__global__ void __launch_bounds__(256) wgrad_kernel(Geom G, const unsigned short *__restrict__ a,
                                                    const unsigned short *__restrict__ dout,
                                                    float *__restrict__ part) {
    const long NP = npix(G);
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const long wave = (long)blockIdx.x * 4 + (threadIdx.x >> 6), nwaves = (long)gridDim.x * 4;
    const long ntiles = (NP + 15) / 16;
    float4m acc[NACC];
#pragma unroll
    for (int t = 0; t < NACC; t++) acc[t] = (float4m){0.f, 0.f, 0.f, 0.f};
    for (long tile = wave; tile < ntiles; tile += nwaves) {
        const long p0 = tile * 16 + 4 * g;
        short4v av, one;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const bool ok = p0 + j < NP;
            av[j] = ok ? (short)dout[(p0 + j) * C + r] : (short)0;
            one[j] = (ok && r == 0) ? (short)0x3f80 : (short)0;   // BF16 1.0 in column 0
        }
#pragma unroll
        for (int t = 0; t < TAPS; t++) {
            short4v bv;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const long q = p0 + j < NP ? nbr(G, p0 + j, t) : -1;
                bv[j] = q >= 0 ? (short)a[q * C + r] : (short)0;
            }
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[t], 0, 0, 0);
        }
        acc[TAPS] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, one, acc[TAPS], 0, 0, 0);
    }
    // D[m = co = 4(l/16)+j][n = ci = l%16]
#pragma unroll
    for (int t = 0; t < NACC; t++)
#pragma unroll
        for (int j = 0; j < 4; j++) part[((wave * NACC + t) * C + 4 * g + j) * C + r] = acc[t][j];
}

// Sum the partials over wavefronts in a fixed order into one layer's slice of
// the parameter gradient ([t][co][ci] weights, then 16 biases).
__global__ void wgrad_reduce_kernel(const float *__restrict__ part, int nwaves, float *__restrict__ grad) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= NACC * C * C) return;
    const int t = idx / (C * C), rem = idx % (C * C);
    if (t == TAPS && rem % C != 0) return;
    float s = 0.f;
    for (int w = 0; w < nwaves; w++) s += part[((long)w * NACC + t) * C * C + rem];
    if (t < TAPS) grad[t * C * C + rem] = s;
    else grad[WL + rem / C] = s;
}

// MSE loss: dy[p] = 2 (y[p] - tgt[p]) / NP in channel 0 (BF16, channels
// 1..15 zero); part[block] = sum (y - tgt)^2 over the block (FP64).
__global__ void __launch_bounds__(256) loss_kernel(long NP, const float *__restrict__ y,
                                                   const float *__restrict__ tgt, unsigned short *__restrict__ dy,
                                                   double *__restrict__ part) {
    __shared__ double sh[256];
    double acc = 0.0;
    const float sc = 2.0f / (float)NP;
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < NP; p += (long)gridDim.x * blockDim.x) {
        const float d = y[p] - tgt[p];
        acc += (double)d * (double)d;
        dy[p * C] = f2bf(sc * d);
#pragma unroll
        for (int c = 1; c < C; c++) dy[p * C + c] = 0;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Adam on the FP32 master parameters; `real` is 1 for trained entries and 0
// for the zero-held padding of the 1-channel layers.
__global__ void adam_kernel(int n, float *__restrict__ theta, const float *__restrict__ grad, float *__restrict__ m,
                            float *__restrict__ v, const float *__restrict__ real, float lr, float b1, float b2,
                            float eps, float bc1, float bc2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gr = grad[i] * real[i];
    const float mi = b1 * m[i] + (1.f - b1) * gr;
    const float vi = b2 * v[i] + (1.f - b2) * gr * gr;
    m[i] = mi;
    v[i] = vi;
    theta[i] = theta[i] - real[i] * lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// FP32 parameters -> BF16 forward weights Wf[l][t][co][ci], backward weights
// Wb[l][t][ci][co] = Wf[l][8-t][co][ci] (taps mirrored, channels swapped: the
// input-gradient convolution), FP32 biases.
__global__ void pack_kernel(const float *__restrict__ theta, unsigned short *__restrict__ Wf,
                            unsigned short *__restrict__ Wb, float *__restrict__ bias) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < LAYERS * WL) {
        const int l = idx / WL, k = idx % WL;
        const int t = k / (C * C), co = (k / C) % C, ci = k % C;
        const unsigned short w = f2bf(theta[l * PL + k]);
        Wf[idx] = w;
        Wb[l * WL + ((TAPS - 1 - t) * C + ci) * C + co] = w;
    }
    if (idx < LAYERS * C) bias[idx] = theta[(idx / C) * PL + WL + idx % C];
}

// Network input from FP32 images: dst[p][0] = bf16(src[p]), other channels 0.
__global__ void input_f32_kernel(long NP, const float *__restrict__ src, unsigned short *__restrict__ dst) {
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < NP; p += (long)gridDim.x * blockDim.x) {
        dst[p * C] = f2bf(src[p]);
#pragma unroll
        for (int c = 1; c < C; c++) dst[p * C + c] = 0;
    }
}

// Network input of the preconditioner: image 0 = q / max|q|, image 1 = -q / max|q|.
__global__ void input_pm_kernel(long P, const double *__restrict__ q, const double *__restrict__ qmax,
                                unsigned short *__restrict__ dst) {
    const double m = *qmax;
    const double s = m > 0.0 ? 1.0 / m : 0.0;
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < P; p += (long)gridDim.x * blockDim.x) {
        const float v = (float)(q[p] * s);
        dst[p * C] = f2bf(v);
        dst[(P + p) * C] = f2bf(-v);
#pragma unroll
        for (int c = 1; c < C; c++) {
            dst[p * C + c] = 0;
            dst[(P + p) * C + c] = 0;
        }
    }
}

// Odd part of the network, rescaled: out[p] = scale * max|q| * (y[p] - y[P+p]) / 2.
__global__ void h0_out_kernel(long P, const float *__restrict__ y, const double *__restrict__ qmax, double scale,
                              double *__restrict__ out) {
    const double m = *qmax * scale * 0.5;
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < P; p += (long)gridDim.x * blockDim.x)
        out[p] = m * ((double)y[p] - (double)y[P + p]);
}

// max |q| (two passes: per-block partials, then one block).
__global__ void maxabs_partial_kernel(long n, const double *__restrict__ q, double *__restrict__ part) {
    __shared__ double sh[256];
    double m = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        m = fmax(m, fabs(q[i]));
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
__global__ void maxabs_final_kernel(int np, const double *__restrict__ part, double *__restrict__ out) {
    __shared__ double sh[256];
    double m = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) m = fmax(m, part[i]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// Training pair k into the image stack: x = g/max|g|, target = -s/max|s| (and
// the negated copy as image K + k), FP32.
__global__ void pair_kernel(long P, const double *__restrict__ gk, const double *__restrict__ gmax,
                            const double *__restrict__ sk, const double *__restrict__ smax, float *__restrict__ x,
                            float *__restrict__ xneg, float *__restrict__ tgt, float *__restrict__ tneg, double sgn) {
    const double a = *gmax > 0.0 ? 1.0 / *gmax : 0.0, b = *smax > 0.0 ? sgn / *smax : 0.0;
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < P; p += (long)gridDim.x * blockDim.x) {
        const float xv = (float)(gk[p] * a), tv = (float)(sk[p] * b);
        x[p] = xv;
        xneg[p] = -xv;
        tgt[p] = tv;
        tneg[p] = -tv;
    }
}

// w = (illum + eps)^(-1/2)
__global__ void illum_weight_kernel(long n, const float *__restrict__ illum, float eps, float *__restrict__ w) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float v = illum[i] + eps;
        w[i] = v > 0.f ? 1.0f / sqrtf(v) : 0.f;
    }
}

// y = s * w * x  (w may be null: 1)
__global__ void scale_w_kernel(long n, const float *__restrict__ w, const double *__restrict__ x, double s,
                               double *__restrict__ y) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = s * (w ? (double)w[i] : 1.0) * x[i];
}

// y = x / w  (w may be null: 1)
__global__ void unscale_w_kernel(long n, const float *__restrict__ w, const double *__restrict__ x,
                                 double *__restrict__ y) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = w ? (w[i] > 0.f ? x[i] / (double)w[i] : 0.0) : x[i];
}

// Coarse grid of the preconditioner (r x r cells of the (nz, nx) column-major
// model per coarse cell): c = the mean of each block (partial blocks at the
// bottom and right edges average the cells they hold).
__global__ void restrict_kernel(const double *__restrict__ f, int nz, int nx, int r, int nzc, int nxc,
                                double *__restrict__ c) {
    const long nc = (long)nzc * nxc;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < nc; q += (long)gridDim.x * blockDim.x) {
        const int I = (int)(q % nzc), J = (int)(q / nzc);
        const int i1 = min((I + 1) * r, nz), j1 = min((J + 1) * r, nx);
        double s = 0.0;
        for (int j = J * r; j < j1; j++)
            for (int i = I * r; i < i1; i++) s += f[i + (long)j * nz];
        c[q] = s / (double)((i1 - I * r) * (j1 - J * r));
    }
}
// Back to the model grid: bilinear between coarse-cell centres ((I + 1/2) r -
// 1/2 in model cells), constant beyond the outermost centres.
__device__ __forceinline__ void lin1(int i, int r, int nc, int &a, int &b, double &w) {
    double x = ((double)i + 0.5) / r - 0.5;
    x = fmin(fmax(x, 0.0), (double)(nc - 1));
    a = (int)x;
    b = min(a + 1, nc - 1);
    w = x - a;
}
__global__ void prolong_kernel(const double *__restrict__ c, int nzc, int nxc, int r, int nz, int nx,
                               double *__restrict__ f) {
    const long n = (long)nz * nx;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        int i0, i1, j0, j1;
        double wi, wj;
        lin1((int)(q % nz), r, nzc, i0, i1, wi);
        lin1((int)(q / nz), r, nxc, j0, j1, wj);
        const double a = (1.0 - wi) * c[i0 + (long)j0 * nzc] + wi * c[i1 + (long)j0 * nzc];
        const double b = (1.0 - wi) * c[i0 + (long)j1 * nzc] + wi * c[i1 + (long)j1 * nzc];
        f[q] = (1.0 - wj) * a + wj * b;
    }
}

// y += x
__global__ void accumulate_kernel(long n, const double *__restrict__ x, double *__restrict__ y) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = y[i] + x[i];
}

}  // namespace cnn
