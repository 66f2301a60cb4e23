// fd_kernels.h — CDNA4 (gfx950) kernels of the acoustic FD engine.
//
// One launch advances every active shot of a batch by one time step.  A
// workgroup owns a TZ x TX tile of one shot's padded grid and performs, in
// the reference's per-step order (libs/libwave/Propagator.f90:168-243):
//   ExtractData (receivers read from u2), ExtractWave (snapshot of u2),
//   ImagingCondition (backward: xcorr of u2 with the source snapshot),
//   FD_scheme (8th-order Laplacian, forward v^2 lap(u) or adjoint-style
//   lap(v^2 u)), Injection (sources/receivers, per-cell contribution lists),
//   Born secondary source, the 5-level time update, and the Higdon absorbing
//   boundary for tiles that touch the grid edge.
// u3 is written over u_{-1}'s buffer: four rotating buffers instead of the
// reference's five (FDswaptimes.f90:14-24), since u_{-1} is only read at the
// point being written.
//
// All arithmetic keeps the reference's operation order and this file is
// compiled with -ffp-contract=off, so FP32 results match the reference's
// kernels to the bit (the oracle tests check it).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

namespace sfd {

constexpr int kThreads = 256;

// Imaging modes (Imaging.f90)
enum { IMG_NONE = 0, IMG_RTM = 1, IMG_FWI = 2, IMG_LSRTM = 3, IMG_BORN = 4 };
// Extraction modes (Extraction.f90), -1 none
enum { EXT_SINC_V2 = 0, EXT_NEAREST = 1, EXT_SINC_LSRTM = 2, EXT_SINC_UNIT = 3 };

// Per-shot device descriptor (all pointers are device pointers).
struct ShotDev {
    float *u[4];            // rotating wavefield buffers, LDZ*NX each
    const float *v2;        // vel*vel, padded
    const float *vel;       // vel, padded (imaging / Born)
    const float *hz;        // Higdon z-band coefficients: 16 per column ix in [4, NX-5]
    const float *hx;        // Higdon x-band coefficients: 16 per row iz in [4, NZ-5]
    int NZ, NX, LDZ, zoff;  // padded dims, leading dim, storage offset of row 0
    int zsh, xsh;           // tile-origin shifts (tile k starts at k*T - sh)
    int ntz, ntx, tile0;    // tile grid and first logical tile
    int nz, nxw, i0;        // window interior and its first padded index (4+ntaper)
    int nt, ntsnap;
    // injection: per tile [inj_tile[t], inj_tile[t+1]) cells; per cell a
    // contribution range [inj_cptr[c], inj_cptr[c+1])
    const int *inj_tile;
    const int *inj_cell;    // local cell index lz + TZ*lx
    const int *inj_cptr;
    const float *inj_a;     // v2r (gain) per contribution
    const float *inj_wz;    // sinc z weight
    const float *inj_wx;    // sinc x weight
    const int *inj_tr;      // trace index per contribution
    const float *inj_traces;// source traces, (nt, ntraces) time fastest
    // extraction: per tile [ext_tile[t], ext_tile[t+1]) receivers
    const int *ext_tile;
    const int *ext_rec;     // receiver index
    const int *ext_pos;     // local centre cell (lz | lx << 16)
    const float *ext_w;     // 16 floats per receiver: gain, wz[7], wx[7], pad
    float *traces;          // extracted traces (nt, ntr)
    int ext_kind;
    // snapshots (window interior, nz*nxw per snapshot)
    void *snap;             // float* or __half*
    float *snap_scale;      // per-snapshot max|S| (FP16 mode)
    unsigned *snap_max;     // running max bits for the next snapshot (FP16 mode)
    // imaging / Born
    float *image, *illum;   // nz*nxw window accumulators
    const float *born_m;    // nz*nxw reflectivity
    const float *taper;     // nxw
    float dt2snap;          // (dt*snapi)^2
};

struct StepArgs {
    const ShotDev *shots;
    const int *blk_shot;    // logical tile -> shot
    int nblocks;            // logical tiles (grid may be padded to a multiple of 8)
    int xcd_remap;          // map block ids so one XCD gets contiguous tiles
    int n;                  // step counter from 0
    int rot;                // n % 4
    int dir;                // +1 forward time, -1 reverse time
    int snapi;
    int store_snap;         // 1: store u2 as the snapshot of this time
    int snap_fp16;          // snapshots are FP16 with a per-snapshot scale
    int imaging;            // IMG_*
    int extract;            // 1: run receiver extraction
    int inject;             // 1: run injection
    float c[10];            // c0z..c4z, c0x..c4x (FDtypes.f90:367-387)
    float c12dt2;           // 12*dt^2
    float div11;            // 1/11
};

__device__ __forceinline__ size_t addr(const ShotDev &S, int iz, int ix) {
    return (size_t)ix * (size_t)S.LDZ + (size_t)(S.zoff + iz);
}

// Logical tile of this workgroup, XCD-aware: consecutive logical tiles (which
// share halos) run on the same XCD and hit its L2.
__device__ __forceinline__ int logical_block(const StepArgs &A) {
    int b = blockIdx.x;
    if (A.xcd_remap) {
        int per = gridDim.x >> 3;
        b = (b & 7) * per + (b >> 3);
    }
    return b;
}

// Snapshot idx (1-based) at window offset wq, decoded to FP32.
__device__ __forceinline__ float load_snap(const ShotDev &S, int fp16, int idx, size_t npix, size_t wq) {
    if (fp16) {
        const __half h = ((const __half *)S.snap)[(size_t)(idx - 1) * npix + wq];
        return __half2float(h) * S.snap_scale[idx - 1];
    }
    return ((const float *)S.snap)[(size_t)(idx - 1) * npix + wq];
}

template <int TZ, int TX, bool ADJ>
__global__ void __launch_bounds__(kThreads) fd_step_kernel(StepArgs A) {
    constexpr int XG = kThreads / TZ;   // x-groups of threads
    constexpr int PX = TX / XG;         // columns per thread
    constexpr int LZ = TZ + 8;          // LDS tile rows incl. halo
    constexpr int LX = TX + 8;
    static_assert(kThreads % TZ == 0 && TX % XG == 0, "tile shape");
    __shared__ float s_a[LX * LZ];      // u2 (forward) or v2*u2 (adjoint), with halo
    __shared__ float s_b[TX * TZ];      // u3 tile for the Higdon recursions
    __shared__ int s_inj[TX * TZ];      // per-cell contribution range (start, count)

    const int lb = logical_block(A);
    if (lb >= A.nblocks) return;
    const int si = A.blk_shot[lb];
    const ShotDev &S = A.shots[si];
    const int it = A.dir > 0 ? A.n + 1 : S.nt - A.n;
    if (it < 1 || it > S.nt) return;   // this shot has finished (ragged batch)

    const int NZ = S.NZ, NX = S.NX;
    const int t = lb - S.tile0;
    const int tzi = t % S.ntz, txi = t / S.ntz;
    const int z0 = tzi * TZ - S.zsh, x0 = txi * TX - S.xsh;
    float *__restrict__ u2 = S.u[(A.rot + 3) & 3];
    const float *__restrict__ u1 = S.u[(A.rot + 2) & 3];
    const float *__restrict__ u0 = S.u[(A.rot + 1) & 3];
    float *__restrict__ u3 = S.u[A.rot];      // u_{-1} in, u3 out
    const int tid = threadIdx.x;

    // ---- stage u2 (or v2(clamped)*u2) tile + 4-cell halo in LDS
    for (int l = tid; l < LX * LZ; l += kThreads) {
        const int lz = l % LZ, lx = l / LZ;
        const int iz = z0 - 4 + lz, ix = x0 - 4 + lx;
        float val = 0.f;
        if (iz >= 0 && iz < NZ && ix >= 0 && ix < NX) {
            val = u2[addr(S, iz, ix)];
            if (ADJ) {
                // FDderivatives.f90:859-876: the velocity index of a neighbour is
                // clamped to the FD region [4, N-5] of each axis.
                const int cz = iz < 4 ? 4 : (iz > NZ - 5 ? NZ - 5 : iz);
                const int cx = ix < 4 ? 4 : (ix > NX - 5 ? NX - 5 : ix);
                val = S.v2[addr(S, cz, cx)] * val;
            }
        }
        s_a[l] = val;
    }

    const int has_inj = A.inject && (S.inj_tile[t + 1] > S.inj_tile[t]);
    if (has_inj) {
        for (int l = tid; l < TX * TZ; l += kThreads) s_inj[l] = 0;
    }
    __syncthreads();
    if (has_inj) {
        for (int k = S.inj_tile[t] + tid; k < S.inj_tile[t + 1]; k += kThreads) {
            const int c0 = S.inj_cptr[k], c1 = S.inj_cptr[k + 1];
            s_inj[S.inj_cell[k]] = (c0 << 8) | (c1 - c0);
        }
    }

    // ---- receivers whose centre lies in this tile (ExtractData, on u2)
    if (A.extract && S.ext_tile) {
        for (int k = S.ext_tile[t] + tid; k < S.ext_tile[t + 1]; k += kThreads) {
            const int r = S.ext_rec[k];
            const int pos = S.ext_pos[k];
            const int lz = (pos & 0xffff) + 4, lx = (pos >> 16) + 4;
            float *tr = S.traces + (size_t)r * S.nt + (it - 1);
            if (S.ext_kind == EXT_NEAREST) {
                // Extraction.f90:84-86
                *tr = *tr + s_a[lz + LZ * lx];
            } else {
                // Extraction.f90:539-545: trace += g*u*wz*wx, j outer, i inner
                const float *w = S.ext_w + 16 * k;
                const float g = w[0];
                float acc = *tr;
#pragma unroll
                for (int j = -3; j <= 3; j++)
#pragma unroll
                    for (int i = -3; i <= 3; i++)
                        acc = acc + g * s_a[(lz + i) + LZ * (lx + j)] * w[4 + i] * w[11 + j];
                *tr = acc;
            }
        }
    }

    const int lane = tid % TZ, xg = tid / TZ;
    const int iz = z0 + lane;
    const bool zin = iz >= 4 && iz <= NZ - 5;
    const bool interior_z = iz >= S.i0 && iz < S.i0 + S.nz;
    const bool snap_step = (it % A.snapi) == 0;
    const int lzc = lane + 4;

    float u3v[PX];
    __syncthreads();   // s_inj complete

#pragma unroll
    for (int p = 0; p < PX; p++) {
        const int lx = xg + p * XG;
        const int ix = x0 + lx;
        const bool xin = ix >= 4 && ix <= NX - 5;
        u3v[p] = 0.f;
        if (!(zin && xin)) continue;
        const size_t q = addr(S, iz, ix);
        const int lxc = lx + 4;
        const float *col = s_a + LZ * lxc;
        const float u2c = ADJ ? u2[q] : col[lzc];

        // snapshot / imaging on u2 (interior window only)
        if (snap_step && interior_z && ix >= S.i0 && ix < S.i0 + S.nxw) {
            const int wz = iz - S.i0, wx = ix - S.i0;
            const size_t wq = (size_t)wx * S.nz + wz;
            const size_t npix = (size_t)S.nz * S.nxw;
            const int idx = it / A.snapi;   // 1-based snapshot index of this time
            if (A.store_snap) {            // Extraction.f90:595
                if (A.snap_fp16) {
                    const float sc = S.snap_scale[idx - 1];
                    const float inv = sc > 0.f ? 1.0f / sc : 0.f;
                    ((__half *)S.snap)[(size_t)(idx - 1) * npix + wq] = __float2half_rn(u2c * inv);
                } else {
                    ((float *)S.snap)[(size_t)(idx - 1) * npix + wq] = u2c;
                }
            }
            if (A.imaging == IMG_RTM || A.imaging == IMG_FWI || A.imaging == IMG_LSRTM) {
                const float Sv = load_snap(S, A.snap_fp16, idx, npix, wq);
                const float tp = S.taper[wx];
                float img = S.image[wq];
                if (A.imaging == IMG_RTM) {
                    img = img + u2c * Sv * tp;                      // Imaging.f90:265
                } else if (A.imaging == IMG_FWI) {
                    const int im1 = idx - 1 > 1 ? idx - 1 : 1;      // Imaging.f90:341-342
                    const int ip1 = idx + 1 < S.ntsnap ? idx + 1 : S.ntsnap;
                    const float Sm = load_snap(S, A.snap_fp16, im1, npix, wq);
                    const float Sp = load_snap(S, A.snap_fp16, ip1, npix, wq);
                    const float v = S.vel[q];
                    const float tmp = (Sm - 2.f * Sv + Sp) / (S.dt2snap * (v * v * v));  // Imaging.f90:365
                    img = img + u2c * tmp * tp;
                } else {
                    img = img + 2.f * tp * u2c * Sv / S.vel[q];     // Imaging.f90:738
                }
                S.image[wq] = img;
                S.illum[wq] = S.illum[wq] + Sv * Sv * tp;
            }
        }

        // FD_scheme: FDderivatives.f90:728-738 (forward) / :877-887 (adjoint)
        const float *c = A.c;
        const float tz = c[0] * (col[lzc]) + c[1] * (col[lzc + 1] + col[lzc - 1]) +
                         c[2] * (col[lzc + 2] + col[lzc - 2]) + c[3] * (col[lzc + 3] + col[lzc - 3]) +
                         c[4] * (col[lzc + 4] + col[lzc - 4]);
        const float tx = c[5] * (col[lzc]) + c[6] * (s_a[lzc + LZ * (lxc + 1)] + s_a[lzc + LZ * (lxc - 1)]) +
                         c[7] * (s_a[lzc + LZ * (lxc + 2)] + s_a[lzc + LZ * (lxc - 2)]) +
                         c[8] * (s_a[lzc + LZ * (lxc + 3)] + s_a[lzc + LZ * (lxc - 3)]) +
                         c[9] * (s_a[lzc + LZ * (lxc + 4)] + s_a[lzc + LZ * (lxc - 4)]);
        float val = ADJ ? (tx + tz) : S.v2[q] * (tx + tz);

        // Injection: per-cell contributions in trace order (Injection.f90:434-440)
        if (has_inj) {
            const int e = s_inj[lane + TZ * lx];
            const int cnt = e & 0xff;
            for (int k = e >> 8, kend = (e >> 8) + cnt; k < kend; k++) {
                const float tr = S.inj_traces[(size_t)S.inj_tr[k] * S.nt + (it - 1)];
                val = val + S.inj_a[k] * tr * S.inj_wz[k] * S.inj_wx[k];
            }
        }
        // Born secondary source: Imaging.f90:779-783, forward time only
        if (A.imaging == IMG_BORN && snap_step && interior_z && ix >= S.i0 && ix < S.i0 + S.nxw) {
            const int wz = iz - S.i0, wx = ix - S.i0;
            const size_t wq = (size_t)wx * S.nz + wz;
            const int idx = it / A.snapi;
            const float Sv = load_snap(S, A.snap_fp16, idx, (size_t)S.nz * S.nxw, wq);
            val = val + 2.f * S.taper[wx] * S.born_m[wq] * Sv / S.vel[q];
        }
        // TimeDer: FDderivatives.f90:974-976
        u3v[p] = (20.f * u2c - 6.f * u1[q] - 4.f * u0[q] + u3[q] + A.c12dt2 * val) * A.div11;
    }

    // ---- Higdon absorbing boundary (Boundary.f90:119-153) for edge tiles
    const bool top = z0 <= 0, bot = z0 + TZ >= NZ, left = x0 <= 0, right = x0 + TX >= NX;
    if (top || bot || left || right) {
        __syncthreads();   // s_inj reads done
#pragma unroll
        for (int p = 0; p < PX; p++) s_b[lane + TZ * (xg + p * XG)] = u3v[p];
        __syncthreads();
        // z-bands: one thread per tile column with ix in [nmin2, nmax2]
        if (top || bot) {
            for (int lx = tid; lx < 2 * TX; lx += kThreads) {
                const int cx = lx % TX;
                const bool dotop = lx < TX;
                if (dotop ? !top : !bot) continue;
                const int ix = x0 + cx;
                if (ix < 4 || ix > NX - 5) continue;
                const float *g = S.hz + 16 * (ix - 4) + (dotop ? 0 : 8);
                float *b = s_b + TZ * cx;
                const float *a = s_a + LZ * (cx + 4) + 4;   // a[lz] = u2 tile (forward)
                if (dotop) {
                    for (int i = 7; i >= 0; i--) {          // nmin1+3 .. nmin1-4
                        const int lz = i - z0;
                        float w0, w1, w2;
                        if (ADJ) {
                            w0 = u2[addr(S, i, ix)]; w1 = u2[addr(S, i + 1, ix)]; w2 = u2[addr(S, i + 2, ix)];
                        } else {
                            w0 = a[lz]; w1 = a[lz + 1]; w2 = a[lz + 2];
                        }
                        const float f = g[0] * b[lz + 1] + g[1] * b[lz + 2] + g[2] * w0 + g[3] * w1 + g[4] * w2 +
                                        g[5] * u1[addr(S, i, ix)] + g[6] * u1[addr(S, i + 1, ix)] +
                                        g[7] * u1[addr(S, i + 2, ix)];
                        b[lz] = -f;
                    }
                } else {
                    for (int i = NZ - 8; i <= NZ - 1; i++) {  // nmax1-3 .. nmax1+4
                        const int lz = i - z0;
                        float w0, w1, w2;
                        if (ADJ) {
                            w0 = u2[addr(S, i, ix)]; w1 = u2[addr(S, i - 1, ix)]; w2 = u2[addr(S, i - 2, ix)];
                        } else {
                            w0 = a[lz]; w1 = a[lz - 1]; w2 = a[lz - 2];
                        }
                        const float f = g[0] * b[lz - 1] + g[1] * b[lz - 2] + g[2] * w0 + g[3] * w1 + g[4] * w2 +
                                        g[5] * u1[addr(S, i, ix)] + g[6] * u1[addr(S, i - 1, ix)] +
                                        g[7] * u1[addr(S, i - 2, ix)];
                        b[lz] = -f;
                    }
                }
            }
            __syncthreads();
        }
        // x-bands: one thread per tile row with iz in [nmin1, nmax1]
        if (left || right) {
            for (int lz = tid; lz < 2 * TZ; lz += kThreads) {
                const int cz = lz % TZ;
                const bool doleft = lz < TZ;
                if (doleft ? !left : !right) continue;
                const int i = z0 + cz;
                if (i < 4 || i > NZ - 5) continue;
                const float *g = S.hx + 16 * (i - 4) + (doleft ? 0 : 8);
                if (doleft) {
                    for (int j = 7; j >= 0; j--) {
                        const int lx = j - x0;
                        float w0, w1, w2;
                        if (ADJ) {
                            w0 = u2[addr(S, i, j)]; w1 = u2[addr(S, i, j + 1)]; w2 = u2[addr(S, i, j + 2)];
                        } else {
                            w0 = s_a[(cz + 4) + LZ * (lx + 4)]; w1 = s_a[(cz + 4) + LZ * (lx + 5)];
                            w2 = s_a[(cz + 4) + LZ * (lx + 6)];
                        }
                        const float f = g[0] * s_b[cz + TZ * (lx + 1)] + g[1] * s_b[cz + TZ * (lx + 2)] +
                                        g[2] * w0 + g[3] * w1 + g[4] * w2 + g[5] * u1[addr(S, i, j)] +
                                        g[6] * u1[addr(S, i, j + 1)] + g[7] * u1[addr(S, i, j + 2)];
                        s_b[cz + TZ * lx] = -f;
                    }
                } else {
                    for (int j = NX - 8; j <= NX - 1; j++) {
                        const int lx = j - x0;
                        float w0, w1, w2;
                        if (ADJ) {
                            w0 = u2[addr(S, i, j)]; w1 = u2[addr(S, i, j - 1)]; w2 = u2[addr(S, i, j - 2)];
                        } else {
                            w0 = s_a[(cz + 4) + LZ * (lx + 4)]; w1 = s_a[(cz + 4) + LZ * (lx + 3)];
                            w2 = s_a[(cz + 4) + LZ * (lx + 2)];
                        }
                        const float f = g[0] * s_b[cz + TZ * (lx - 1)] + g[1] * s_b[cz + TZ * (lx - 2)] +
                                        g[2] * w0 + g[3] * w1 + g[4] * w2 + g[5] * u1[addr(S, i, j)] +
                                        g[6] * u1[addr(S, i, j - 1)] + g[7] * u1[addr(S, i, j - 2)];
                        s_b[cz + TZ * lx] = -f;
                    }
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int p = 0; p < PX; p++) u3v[p] = s_b[lane + TZ * (xg + p * XG)];
    }

    // ---- write u3: every non-corner cell of the padded grid that lies in this
    // tile (FD region from the time update, edge bands from Higdon).  The max
    // of |u3| over the window feeds the next snapshot's FP16 scale.
    const bool prev_of_snap = A.store_snap && A.snap_fp16 &&
        ((A.dir > 0 ? it + 1 : it - 1) % A.snapi) == 0;
    float vmax = 0.f;
#pragma unroll
    for (int p = 0; p < PX; p++) {
        const int ix = x0 + xg + p * XG;
        if (iz < 0 || iz >= NZ || ix < 0 || ix >= NX) continue;
        const bool zband = iz >= 4 && iz <= NZ - 5, xband = ix >= 4 && ix <= NX - 5;
        if (!zband && !xband) continue;   // corner: never written by the reference
        u3[addr(S, iz, ix)] = u3v[p];
        if (prev_of_snap && interior_z && ix >= S.i0 && ix < S.i0 + S.nxw) vmax = fmaxf(vmax, fabsf(u3v[p]));
    }
    if (prev_of_snap) {
        // wavefront max, then one atomic per wavefront (non-negative floats order as uints)
        for (int off = 32; off > 0; off >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, off));
        if ((tid & 63) == 0 && vmax > 0.f) atomicMax(S.snap_max, __float_as_uint(vmax));
    }
}

// Edge-replicating window + pad of the full model for one shot
// (mod_window_pad ExtractPadModel.f90:509-530 + model_pad ReadSouVelRho.f90:545-596).
// colsrc[j-1] = model column (0-based) feeding window column j.
__global__ void window_pad_kernel(const float *__restrict__ big, int nz, const int *__restrict__ colsrc,
                                  int nxw, int ntaper, int NZ, int NX, int LDZ, int zoff,
                                  float *__restrict__ vel, float *__restrict__ v2) {
    const int iz = blockIdx.x * blockDim.x + threadIdx.x;
    const int ix = blockIdx.y;
    if (iz >= NZ || ix >= NX) return;
    const int i0 = 4 + ntaper;
    int wz = iz - i0, wx = ix - i0;
    wz = wz < 0 ? 0 : (wz >= nz ? nz - 1 : wz);
    wx = wx < 0 ? 0 : (wx >= nxw ? nxw - 1 : wx);
    const float v = (float)0 + big[(size_t)colsrc[wx] * nz + wz];
    const size_t q = (size_t)ix * LDZ + zoff + iz;
    vel[q] = v;
    v2[q] = v * v;
}

// Boundary.f90:66-87 Higdon_coeff
__device__ __forceinline__ void higdon_coeff(float nu, float *c) {
    const float beta1 = 1.0f, beta2 = 0.7f, weight = 0.5f;
    float fdum = (beta1 + nu) * (1.f - weight);
    const float qx = (weight * (beta1 + nu) - nu) / fdum;
    const float qt = (weight * (beta1 + nu) - beta1) / fdum;
    const float qxt = weight / (weight - 1.f);
    fdum = (beta2 + nu) * (1.f - weight);
    const float rx = (weight * (beta2 + nu) - nu) / fdum;
    const float rt = (weight * (beta2 + nu) - beta2) / fdum;
    const float rxt = weight / (weight - 1.f);
    c[0] = qx + rx;
    c[1] = qx * rx;
    c[2] = qt + rt;
    c[3] = qx * rt + qt * rx + qxt + rxt;
    c[4] = qx * rxt + rx * qxt;
    c[5] = qt * rt;
    c[6] = qt * rxt + rt * qxt;
    c[7] = qxt * rxt;
}

// Boundary.f90:36-50 Higdon tables from the padded velocity at the FD edges
__global__ void higdon_tables_kernel(const float *__restrict__ vel, int NZ, int NX, int LDZ, int zoff,
                                     float dt, float dz, float dx, float *__restrict__ hz,
                                     float *__restrict__ hx) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const float dzi = 1.f / dz, dxi = 1.f / dx;
    if (k < NX - 8) {
        const int ix = k + 4;
        float nu = vel[(size_t)ix * LDZ + zoff + 4] * dt * dzi;
        higdon_coeff(nu, hz + 16 * k);
        nu = vel[(size_t)ix * LDZ + zoff + NZ - 5] * dt * dzi;
        higdon_coeff(nu, hz + 16 * k + 8);
    }
    if (k < NZ - 8) {
        const int iz = k + 4;
        float nu = vel[(size_t)4 * LDZ + zoff + iz] * dt * dxi;
        higdon_coeff(nu, hx + 16 * k);
        nu = vel[(size_t)(NX - 5) * LDZ + zoff + iz] * dt * dxi;
        higdon_coeff(nu, hx + 16 * k + 8);
    }
}

// Publish each shot's running max as the scale of snapshot idx and reset it.
__global__ void snap_scale_kernel(const ShotDev *shots, int nshots, int idx) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshots) return;
    const ShotDev &S = shots[s];
    if (idx >= 1 && idx <= S.ntsnap) S.snap_scale[idx - 1] = __uint_as_float(*S.snap_max);
    *S.snap_max = 0u;
}

// mod_copy_image (ExtractPadModel.f90:662-695): nearest-index scatter-add of
// every shot's window image into the full model.  One thread per model cell;
// shots are summed in shot order, so the result is deterministic.
struct CopyImg {
    const float *img;       // nz*nxw
    const int *colmap;      // window column -> model column or -1
    int nxw;
};
__global__ void copy_image_kernel(const CopyImg *__restrict__ shots, int nshots, int nz, int nx,
                                  const int *__restrict__ inv_ptr, const int *__restrict__ inv_idx,
                                  float *__restrict__ out) {
    // inv_ptr/inv_idx: for model column ix, the (shot, window column) pairs
    // that map to it, in shot order, packed as shot<<20 | wcol.
    const int iz = blockIdx.x * blockDim.x + threadIdx.x;
    const int ix = blockIdx.y;
    if (iz >= nz || ix >= nx) return;
    float acc = out[(size_t)ix * nz + iz];
    for (int k = inv_ptr[ix]; k < inv_ptr[ix + 1]; k++) {
        const int e = inv_idx[k];
        const CopyImg &c = shots[e >> 20];
        acc = acc + c.img[(size_t)(e & 0xfffff) * nz + iz];
    }
    out[(size_t)ix * nz + iz] = acc;
}

}  // namespace sfd
