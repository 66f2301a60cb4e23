// bandpass_kernels.h — Butterworth band-pass of traces (libs/libwave/bandpass.c
// highcut_/lowcut_ inside Bandpass.f90:89-147 bandpass) on the device.
//
// The filter is recursive in time, so one thread runs one trace.  The traces
// are first transposed into a time-major work array (sample i of trace k at
// i * ld + k), which makes every step of the recursion one coalesced load and
// store across a wavefront, padded by 10 samples each side with sin^2 tapers
// of the end values.  A run of the reference applies its sections one after
// another over the whole record (newdata <- section(data); data <- newdata);
// here all sections of one direction run in one sweep, each section taking the
// previous one's output of the same sample, which is the same arithmetic in
// the same order: results are the reference's bit for bit.
#pragma once
#include <hip/hip_runtime.h>

namespace sfd {

constexpr int kBpPad = 10;       // npad (Bandpass.f90:107)
constexpr int kBpMaxSections = 8;

// (ntr, nt) trace-major -> (nt + 2 npad, ld) time-major with the end tapers
// (Bandpass.f90:123-128); 64 x 64 tiles through LDS.
__global__ void __launch_bounds__(256) bp_pad_kernel(const float *__restrict__ in, int nt, int ntr, int ld,
                                                     const float *__restrict__ taper_lo,
                                                     const float *__restrict__ taper_hi, float *__restrict__ work) {
    __shared__ float tile[64][65];
    const int k0 = blockIdx.y * 64, i0 = blockIdx.x * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < 64; r += 4) {   // rows of the tile are traces; coalesced along time
        const int k = k0 + r, i = i0 + tx;
        tile[r][tx] = (k < ntr && i < nt) ? in[(size_t)k * nt + i] : 0.f;
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {   // rows are samples now; coalesced along traces
        const int i = i0 + r, k = k0 + tx;
        if (i < nt && k < ntr) work[(size_t)(i + kBpPad) * ld + k] = tile[tx][r];
    }
    if (blockIdx.x == 0) {               // the pads, from each trace's first and last sample
        for (int j = ty; j < kBpPad; j += 4) {
            const int k = k0 + tx;
            if (k < ntr) {
                work[(size_t)j * ld + k] = in[(size_t)k * nt] * taper_lo[j];
                work[(size_t)(nt + kBpPad + j) * ld + k] = in[(size_t)k * nt + nt - 1] * taper_hi[j];
            }
        }
    }
}

// work rows [npad, npad + nt) back to (ntr, nt) trace-major
__global__ void __launch_bounds__(256) bp_unpad_kernel(const float *__restrict__ work, int nt, int ntr, int ld,
                                                       float *__restrict__ out) {
    __shared__ float tile[64][65];
    const int k0 = blockIdx.y * 64, i0 = blockIdx.x * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < 64; r += 4) {
        const int i = i0 + r, k = k0 + tx;
        tile[r][tx] = (i < nt && k < ntr) ? work[(size_t)(i + kBpPad) * ld + k] : 0.f;
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {
        const int k = k0 + r, i = i0 + tx;
        if (k < ntr && i < nt) out[(size_t)k * nt + i] = tile[tx][r];
    }
}

// All NS sections over one direction of the record.  Forward: samples 2..n-1
// in order, the first section's history primed with samples 0 and 1, which
// the reference then overwrites with the (never written) newdata(0:1) = 0.
// Reverse (bandpass.c:124-138, :249-263): samples n-1 down to 2 with zero
// history (tempdata(0:1) = 0), samples 0 and 1 untouched.
template <int NS>
__device__ __forceinline__ void bp_sweep(float *x, int ld, int n, const float *c, bool reverse) {
    float in1[NS], in2[NS], o1[NS], o2[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) in1[s] = in2[s] = o1[s] = o2[s] = 0.f;
    if (!reverse) {
        in2[0] = x[0];
        in1[0] = x[(size_t)ld];
    }
    constexpr int U = 16;   // samples loaded ahead of the recursion
    for (int i0 = 2; i0 < n; i0 += U) {
        float v[U];
        const int m = n - i0 < U ? n - i0 : U;
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = i0 + u, p = reverse ? n + 1 - i : i;
            if (u < m) v[u] = x[(size_t)p * ld];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (u < m) {
                float a = v[u];
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    const float *b = c + 5 * s;
                    const float y = b[0] * a + b[1] * in1[s] + b[2] * in2[s] - b[3] * o1[s] - b[4] * o2[s];
                    in2[s] = in1[s];
                    in1[s] = a;
                    o2[s] = o1[s];
                    o1[s] = y;
                    a = y;
                }
                v[u] = a;
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = i0 + u, p = reverse ? n + 1 - i : i;
            if (u < m) x[(size_t)p * ld] = v[u];
        }
    }
    if (!reverse) {
        x[0] = 0.f;
        x[(size_t)ld] = 0.f;
    }
}

template <int NS>
__device__ void bp_cascade(float *x, int ld, int n, const float *c, int phase) {
    bp_sweep<NS>(x, ld, n, c, false);
    if (phase == 0) bp_sweep<NS>(x, ld, n, c, true);
}

__device__ __forceinline__ void bp_dispatch(float *x, int ld, int n, const float *c, int ns, int phase) {
    switch (ns) {
        case 1: bp_cascade<1>(x, ld, n, c, phase); break;
        case 2: bp_cascade<2>(x, ld, n, c, phase); break;
        case 3: bp_cascade<3>(x, ld, n, c, phase); break;
        case 4: bp_cascade<4>(x, ld, n, c, phase); break;
        case 5: bp_cascade<5>(x, ld, n, c, phase); break;
        case 6: bp_cascade<6>(x, ld, n, c, phase); break;
        case 7: bp_cascade<7>(x, ld, n, c, phase); break;
        default: bp_cascade<8>(x, ld, n, c, phase); break;
    }
}

// One thread per trace: low-cut then high-cut (Bandpass.f90:130-136), each
// forward and, for phase 0, reversed.  nlo / nhi = 0 skips that filter.
__global__ void __launch_bounds__(64) bp_iir_kernel(float *__restrict__ work, int ld, int ntr, int n,
                                                    const float *__restrict__ coefs, int nlo, int nhi, int phase) {
    __shared__ float c[2 * kBpMaxSections * 5];
    for (int q = threadIdx.x; q < 2 * kBpMaxSections * 5; q += blockDim.x) c[q] = coefs[q];
    __syncthreads();
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ntr) return;
    float *x = work + k;
    if (nlo > 0) bp_dispatch(x, ld, n, c, nlo, phase);
    if (nhi > 0) bp_dispatch(x, ld, n, c + kBpMaxSections * 5, nhi, phase);
}

}  // namespace sfd
