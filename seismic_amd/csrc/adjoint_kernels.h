// adjoint_kernels.h — exact discrete adjoint of the forward modeling time
// step (no reference counterpart: the reference's reverse-time pass uses the
// adjoint-style stencil and re-applies the absorbing boundary, which is not
// the transpose of the forward operator).
//
// Forward step it = k maps (U_k, U_{k-1}, U_{k-2}, U_{k-3}) and the source
// sample s(k) to U_{k+1}: stencil + injection, 5-level update on the FD
// region, then the Higdon bands as sequential overwrites (z-bands, then
// x-bands); extraction reads U_k.  Backwards, with L_j the cotangent of U_j:
//   k7_gamma_kernel  : G = L_{k+1} + E^T dbar(k+1); the Higdon overwrites in
//                      reverse order (x-bands, then z-bands), each
//                      "a = G(c); G(c) = 0; inputs += -g a", adding the u2 / u1
//                      terms into L_k / L_{k-1}; W = 12 dt^2/11 v^2 G (FD cells)
//   k7_update_kernel : L_k += 20/11 G + stencil^T W, L_{k-1} -= 6/11 G,
//                      L_{k-2} -= 4/11 G, L_{k-3} = G/11 (its first term)
//   k7_source_kernel : sbar(k) = 12 dt^2/11 sum over the source footprint of
//                      v2 wz wx G
#pragma once
#include <hip/hip_runtime.h>

namespace sfd {

struct AdjStep {
    ShotDev S;          // geometry, v2, Higdon tables; inj_* = receivers with their extraction
                        // weights, inj_traces = dbar (nt, ntr)
    float *G, *L0, *L1, *L2, *L3, *W;   // L_{k+1} (turned into G in place), L_k .. L_{k-3}, W
    int k;              // forward step whose transpose is applied
    int nsrc;           // source footprint entries
    const int *src_addr;
    const float *src_coef;
    float *sbar;
    float c[10];
    float c12dt2, div11;
};

template <int TZ, int TX>
__global__ void __launch_bounds__(256) k7_gamma_kernel(AdjStep A) {
    const ShotDev &S = A.S;
    __shared__ float s_g[TZ * TX];
    __shared__ int s_inj[TZ * TX];
    __shared__ float s_trv[kMaxTileTraces];
    const int t = blockIdx.x;
    const int tzi = t % S.ntz, txi = t / S.ntz;
    const int z0 = tzi * TZ - S.zsh, x0 = txi * TX - S.xsh;
    const int NZ = S.NZ, NX = S.NX, LDZ = S.LDZ, zoff = S.zoff;
    const int tid = threadIdx.x, NT = blockDim.x;
    for (int l = tid; l < TZ * TX; l += NT) {
        const int iz = z0 + l % TZ, ix = x0 + l / TZ;
        s_g[l] = (iz >= 0 && iz < NZ && ix >= 0 && ix < NX) ? A.G[(size_t)ix * LDZ + zoff + iz] : 0.f;
    }
    // transpose of the extraction at step k+1 (receivers as sources, extraction weights)
    const int it = A.k + 1;
    const int *inj_tile = S.inj_tile;
    if (it <= S.nt && inj_tile[t + 1] > inj_tile[t]) {
        for (int l = tid; l < TZ * TX; l += NT) s_inj[l] = 0;
        const int j0 = S.inj_tptr[t], j1 = S.inj_tptr[t + 1];
        for (int j = j0 + tid; j < j1; j += NT) s_trv[j - j0] = S.inj_traces[(size_t)S.inj_ttr[j] * S.nt + (it - 1)];
        __syncthreads();
        for (int q = inj_tile[t] + tid; q < inj_tile[t + 1]; q += NT) {
            const int c0 = S.inj_cptr[q], c1 = S.inj_cptr[q + 1];
            s_inj[S.inj_cell[q]] = (c0 << 8) | (c1 - c0);
        }
        __syncthreads();
        for (int l = tid; l < TZ * TX; l += NT) {
            const int e = s_inj[l];
            float v = s_g[l];
            for (int q = e >> 8, qe = (e >> 8) + (e & 0xff); q < qe; q++) {
                const InjCoef w = S.inj_coef[q];
                v = v + w.a * s_trv[w.tr] * w.wz * w.wx;
            }
            s_g[l] = v;
        }
    }
    __syncthreads();
    const bool top = z0 <= 0, bot = z0 + TZ >= NZ, left = x0 <= 0, right = x0 + TX >= NX;
#define LOC(i, j) ((i) - z0 + TZ * ((j) - x0))
#define ADR(i, j) ((size_t)(j) * LDZ + zoff + (i))
    // one reversed overwrite u3(i,j) <- -(g0 u3(+1) + g1 u3(+2) + g2 u2 + g3 u2(+1) + g4 u2(+2)
    //                                      + g5 u1 + g6 u1(+1) + g7 u1(+2)), "+1" = one cell along (di, dj)
    auto adj_op = [&](const float *g, int i, int j, int di, int dj) {
        const float a = s_g[LOC(i, j)];
        s_g[LOC(i, j)] = 0.f;
        s_g[LOC(i + di, j + dj)] = s_g[LOC(i + di, j + dj)] - g[0] * a;
        s_g[LOC(i + 2 * di, j + 2 * dj)] = s_g[LOC(i + 2 * di, j + 2 * dj)] - g[1] * a;
        float *L0 = A.L0, *L1 = A.L1;
        L0[ADR(i, j)] = L0[ADR(i, j)] - g[2] * a;
        L0[ADR(i + di, j + dj)] = L0[ADR(i + di, j + dj)] - g[3] * a;
        L0[ADR(i + 2 * di, j + 2 * dj)] = L0[ADR(i + 2 * di, j + 2 * dj)] - g[4] * a;
        L1[ADR(i, j)] = L1[ADR(i, j)] - g[5] * a;
        L1[ADR(i + di, j + dj)] = L1[ADR(i + di, j + dj)] - g[6] * a;
        L1[ADR(i + 2 * di, j + 2 * dj)] = L1[ADR(i + 2 * di, j + 2 * dj)] - g[7] * a;
    };
    if (left || right) {   // x-bands of rows [4, NZ-5], reversed
        for (int l = tid; l < 2 * TZ; l += NT) {
            const bool dl = l < TZ;
            if (dl ? !left : !right) continue;
            const int i = z0 + l % TZ;
            if (i < 4 || i > NZ - 5) continue;
            const float *g = S.hx + 16 * (i - 4) + (dl ? 0 : 8);
            if (dl)
                for (int j = 0; j <= 7; j++) adj_op(g, i, j, 0, 1);
            else
                for (int j = NX - 1; j >= NX - 8; j--) adj_op(g, i, j, 0, -1);
        }
        __syncthreads();
    }
    if (top || bot) {      // z-bands of columns [4, NX-5], reversed
        for (int l = tid; l < 2 * TX; l += NT) {
            const bool dtop = l < TX;
            if (dtop ? !top : !bot) continue;
            const int ix = x0 + l % TX;
            if (ix < 4 || ix > NX - 5) continue;
            const float *g = S.hz + 16 * (ix - 4) + (dtop ? 0 : 8);
            if (dtop)
                for (int i = 0; i <= 7; i++) adj_op(g, i, ix, 1, 0);
            else
                for (int i = NZ - 1; i >= NZ - 8; i--) adj_op(g, i, ix, -1, 0);
        }
        __syncthreads();
    }
    for (int l = tid; l < TZ * TX; l += NT) {
        const int iz = z0 + l % TZ, ix = x0 + l / TZ;
        if (iz < 0 || iz >= NZ || ix < 0 || ix >= NX) continue;
        const size_t q = ADR(iz, ix);
        const bool fd = iz >= 4 && iz <= NZ - 5 && ix >= 4 && ix <= NX - 5;
        A.G[q] = s_g[l];
        A.W[q] = fd ? A.c12dt2 * A.div11 * S.v2[q] * s_g[l] : 0.f;
    }
#undef LOC
#undef ADR
}

__global__ void __launch_bounds__(256) k7_update_kernel(AdjStep A) {
    const ShotDev &S = A.S;
    const int NZ = S.NZ, NX = S.NX, LDZ = S.LDZ, zoff = S.zoff;
    const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int iz = (int)(idx % NZ), ix = (int)(idx / NZ);
    if (ix >= NX) return;
    const bool zin = iz >= 4 && iz <= NZ - 5, xin = ix >= 4 && ix <= NX - 5;
    if (!zin && !xin) return;   // corner cells: never read or written
    const size_t q = (size_t)ix * LDZ + zoff + iz;
    const float *W = A.W, *c = A.c;
    auto w = [&](int i, int j) { return (i >= 0 && i < NZ && j >= 0 && j < NX) ? W[(size_t)j * LDZ + zoff + i] : 0.f; };
    float st = (c[0] + c[5]) * W[q];
    for (int t = 1; t <= 4; t++)
        st = st + c[t] * (w(iz + t, ix) + w(iz - t, ix)) + c[5 + t] * (w(iz, ix + t) + w(iz, ix - t));
    const float gq = (zin && xin) ? A.G[q] : 0.f;
    A.L0[q] = A.L0[q] + 20.f * A.div11 * gq + st;
    A.L1[q] = A.L1[q] - 6.f * A.div11 * gq;
    A.L2[q] = A.L2[q] - 4.f * A.div11 * gq;
    A.L3[q] = A.div11 * gq;
}

__global__ void k7_source_kernel(AdjStep A) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float acc = 0.f;
    for (int j = 0; j < A.nsrc; j++) acc = acc + A.src_coef[j] * A.G[A.src_addr[j]];
    A.sbar[A.k - 1] = A.c12dt2 * A.div11 * acc;
}

// Transpose of Injection_Born (Imaging.f90:779-783) at step k: the forward
// adds 2 taper m S_k / v to the stencil value of every window cell, which the
// time update scales by c12dt2 div11, so mbar(window) += that factor times G.
// S.snap holds the FP32 snapshots of the background pass (index k / snapi).
__global__ void __launch_bounds__(256) k7_born_kernel(AdjStep A, int snap_idx, float *mbar) {
    const ShotDev &S = A.S;
    const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (idx >= (long)S.nz * S.nxw) return;
    const int wz = (int)(idx % S.nz), wx = (int)(idx / S.nz);
    const int iz = wz + S.iz0, ix = wx + S.i0;
    if (iz < 4 || iz > S.NZ - 5 || ix < 4 || ix > S.NX - 5) return;
    const size_t q = (size_t)ix * S.LDZ + S.zoff + iz;
    const float Sv = ((const float *)S.snap)[(size_t)(snap_idx - 1) * S.nz * S.nxw + idx];
    const float b = 2.f * S.taper[wx] * Sv / S.vel[q];
    mbar[idx] = mbar[idx] + A.c12dt2 * A.div11 * A.G[q] * b;
}

}  // namespace sfd
