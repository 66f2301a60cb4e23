// fd_step2.h — time blocking: two steps of the time loop in one launch.
//
// A workgroup owns the one-step kernel's 64x32 tile.  It reads U_n on the
// tile plus an 8-cell ring into LDS, computes U_{n+1} on the tile plus a
// 4-cell ring (stencil, injection, 5-level update, then the Higdon bands of
// every edge line that crosses that region), writes U_{n+1} on the tile, and
// computes U_{n+2} on the tile from the LDS copy.  Per cell the arithmetic is
// the one-step kernel's in the same order, so the results are bitwise equal;
// the ring is recomputed by the neighbouring workgroups (1.4x the stencil
// work) while U_n..U_{n-3} and v^2 cross HBM once for two steps and the
// launch count halves.  Extraction, snapshots and imaging of both steps act
// on the tile's own cells.  The Born source term is not handled here (the
// one-step kernel runs Born passes).
//
// Injection uses a per-tile plan over the 72x40 region (the tile's own cells
// and the ring cells of its neighbours' plans), built on the host.
#pragma once
#include <hip/hip_runtime.h>

namespace sfd {

constexpr int B2Z = 64, B2X = 32;               // own tile
constexpr int R1Z = B2Z + 8, R1X = B2X + 8;     // U_{n+1} region (ring 4)
constexpr int R0Z = B2Z + 16, R0X = B2X + 16;   // U_n region (ring 8)
constexpr int kMaxTileTraces2 = 512;            // distinct traces one region may inject
constexpr int kCellAhead = 8;                   // contributions per cell loaded before use
constexpr int kThreads2 = 512;

// One cell's injection, contributions c0 .. c0+cnt-1 in order (cnt >= 1); the
// first kCellAhead coefficient loads are issued before any is used.
__device__ __forceinline__ float inject_cell(float val, const GLB float4v *co, int c0, int cnt, const float *trv) {
    float4v w[kCellAhead];
#pragma unroll
    for (int j = 0; j < kCellAhead; j++) w[j] = co[c0 + (j < cnt ? j : cnt - 1)];
#pragma unroll
    for (int j = 0; j < kCellAhead; j++)
        if (j < cnt) val = val + w[j][0] * trv[__float_as_int(w[j][3])] * w[j][1] * w[j][2];
    for (int j = kCellAhead; j < cnt; j++) {
        const float4v q = co[c0 + j];
        val = val + q[0] * trv[__float_as_int(q[3])] * q[1] * q[2];
    }
    return val;
}

// Receivers centred in the tile, footprint read from an LDS copy of the
// wavefield (origin oz, ox, column stride sz); same summation order as
// extract_tile.
__device__ __forceinline__ void extract_lds(const ShotDev &S, int t, int it, const float *u, int oz, int ox, int sz,
                                            int z0, int x0) {
    const GLB int *et = gp(S.ext_tile);
    for (int k = et[t] + (int)threadIdx.x; k < et[t + 1]; k += (int)blockDim.x) {
        const int r = gp(S.ext_rec)[k];
        const int pos = gp(S.ext_pos)[k];
        const int iz = z0 + (pos & 0xffff), ix = x0 + (pos >> 16);
        GLB float *tr = gp(S.traces) + (size_t)r * S.nt + (it - 1);
#define UL(i, j) u[((i) - oz) + sz * ((j) - ox)]
        if (S.ext_kind == EXT_NEAREST) {
            *tr = *tr + UL(iz, ix);
        } else {
            const GLB float *w = gp(S.ext_w) + 16 * k;
            const float g = w[0];
            float acc = *tr;
            for (int j = -3; j <= 3; j++)
                for (int i = -3; i <= 3; i++) acc = acc + g * UL(iz + i, ix + j) * w[4 + i] * w[11 + j];
            *tr = acc;
        }
#undef UL
    }
}

// Snapshot / imaging of one step on the tile's own cells; u(lz, lx) gives u2.
template <typename U>
__device__ __forceinline__ void snap_image_tile(const ShotDev &S, const StepArgs &A, int t, int it, int z0, int x0,
                                                float *s_red, U u) {
    const int tid = threadIdx.x;
    // FP16 scales per 64 x 4 block: column maxima by wavefront reductions
    // (a wavefront's 64 lanes walk one tile column), then the four columns
    // of each block; s_red holds B2X column maxima and B2X / 4 block scales.
    if (A.store_snap && A.snap_fp16) {
        for (int l = tid; l < B2Z * B2X; l += kThreads2) {
            const int iz = z0 + l % B2Z, ix = x0 + l / B2Z;
            float m = 0.f;
            if (iz >= S.iz0 && iz < S.iz0 + S.nz && ix >= S.i0 && ix < S.i0 + S.nxw) m = fabsf(u(l % B2Z, l / B2Z));
            for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
            if ((tid & 63) == 0) s_red[l / B2Z] = m;
        }
        __syncthreads();
        if (tid < B2X / 4) {
            float m = fmaxf(fmaxf(s_red[4 * tid], s_red[4 * tid + 1]), fmaxf(s_red[4 * tid + 2], s_red[4 * tid + 3]));
            if (!(m > 1e-30f)) m = 0.f;
            gp(S.snap_scale)[(size_t)(it / A.snapi - 1) * S.nsb + snap_block(S, z0, x0 + 4 * tid)] = m;
            s_red[B2X + tid] = m;
        }
        __syncthreads();
    }
    const bool img = A.imaging == IMG_RTM || A.imaging == IMG_FWI || A.imaging == IMG_LSRTM;
    const size_t npix = (size_t)S.nz * S.nxw;
    for (int l = tid; l < B2Z * B2X; l += kThreads2) {
        const int iz = z0 + l % B2Z, ix = x0 + l / B2Z;
        if (!(iz >= S.iz0 && iz < S.iz0 + S.nz && ix >= S.i0 && ix < S.i0 + S.nxw)) continue;
        float Sv = 0.f, img0 = 0.f, ill0 = 0.f, tp = 0.f, vel = 0.f;
        const int sb = snap_block(S, iz, ix);
        float inv16 = 0.f;
        if (A.store_snap && A.snap_fp16) {
            const float m = s_red[B2X + (l / B2Z) / 4];
            inv16 = m > 0.f ? 1.0f / m : 0.f;
        }
        if (img) {
            const size_t wq = (size_t)(ix - S.i0) * S.nz + (iz - S.iz0);
            Sv = load_snap(S, A.snap_fp16, it / A.snapi, npix, wq, sb, true);
            img0 = gp(S.image)[wq];
            ill0 = gp(S.illum)[wq];
            tp = gp(S.taper)[ix - S.i0];
            if (A.imaging == IMG_FWI || A.imaging == IMG_LSRTM) vel = gp(S.vel)[(size_t)ix * S.LDZ + S.zoff + iz];
        }
        if (A.imaging == IMG_FWI)
            snap_image_point<true>(S, A, it, iz, ix, u(l % B2Z, l / B2Z), vel, sb, inv16, Sv, img0, ill0, tp);
        else
            snap_image_point<false>(S, A, it, iz, ix, u(l % B2Z, l / B2Z), vel, sb, inv16, Sv, img0, ill0, tp);
    }
}

// Higdon bands (Boundary.f90:119-153) of the lines crossing rows [r0, r1] x
// columns [c0, c1]: z-bands, then x-bands, each line a sequential recursion in
// the reference's order, computed only on the part of the line inside the
// region (its cells read the line's next two cells, which lie inside too).
// U3 / U2 / U1 give the new level, u2 and u1.
template <typename F3, typename F2, typename F1>
__device__ __forceinline__ void higdon_region(const ShotDev &S, int r0, int r1, int c0, int c1, F3 U3, F2 U2,
                                              F1 U1) {
    const int NZ = S.NZ, NX = S.NX, tid = threadIdx.x;
    const int nc = c1 - c0 + 1, nr = r1 - r0 + 1;
    const bool top = r0 <= 7, bot = r1 >= NZ - 8, left = c0 <= 7, right = c1 >= NX - 8;
    // a line's u2 / u1 over its 10 cells are read in one burst before the
    // recursion (band cell k of the line: k = 0..9 from the edge inwards)
    if (top || bot) {
        for (int l = tid; l < 2 * nc; l += kThreads2) {
            const bool dt = l < nc;
            if (dt ? !top : !bot) continue;
            const int ix = c0 + l % nc;
            if (ix < 4 || ix > NX - 5) continue;
            const GLB float *g = gp(S.hz) + 16 * (ix - 4) + (dt ? 0 : 8);
            float a2[10], a1[10];
#pragma unroll
            for (int k = 0; k < 10; k++) {   // only the cells of the line inside the region
                const int i = dt ? k : NZ - 1 - k;
                const bool in = dt ? i >= r0 : i <= r1;
                a2[k] = in ? U2(i, ix) : 0.f;
                a1[k] = in ? U1(i, ix) : 0.f;
            }
            if (dt) {
#pragma unroll
                for (int i = 7; i >= 0; i--) {
                    if (i < r0) break;
                    const float f = g[0] * U3(i + 1, ix) + g[1] * U3(i + 2, ix) + g[2] * a2[i] + g[3] * a2[i + 1] +
                                    g[4] * a2[i + 2] + g[5] * a1[i] + g[6] * a1[i + 1] + g[7] * a1[i + 2];
                    U3(i, ix) = -f;
                }
            } else {
#pragma unroll
                for (int k = 7; k >= 0; k--) {
                    const int i = NZ - 1 - k;
                    if (i > r1) break;
                    const float f = g[0] * U3(i - 1, ix) + g[1] * U3(i - 2, ix) + g[2] * a2[k] + g[3] * a2[k + 1] +
                                    g[4] * a2[k + 2] + g[5] * a1[k] + g[6] * a1[k + 1] + g[7] * a1[k + 2];
                    U3(i, ix) = -f;
                }
            }
        }
        __syncthreads();
    }
    if (left || right) {
        for (int l = tid; l < 2 * nr; l += kThreads2) {
            const bool dl = l < nr;
            if (dl ? !left : !right) continue;
            const int i = r0 + l % nr;
            if (i < 4 || i > NZ - 5) continue;
            const GLB float *g = gp(S.hx) + 16 * (i - 4) + (dl ? 0 : 8);
            float a2[10], a1[10];
#pragma unroll
            for (int k = 0; k < 10; k++) {
                const int j = dl ? k : NX - 1 - k;
                const bool in = dl ? j >= c0 : j <= c1;
                a2[k] = in ? U2(i, j) : 0.f;
                a1[k] = in ? U1(i, j) : 0.f;
            }
            if (dl) {
#pragma unroll
                for (int j = 7; j >= 0; j--) {
                    if (j < c0) break;
                    const float f = g[0] * U3(i, j + 1) + g[1] * U3(i, j + 2) + g[2] * a2[j] + g[3] * a2[j + 1] +
                                    g[4] * a2[j + 2] + g[5] * a1[j] + g[6] * a1[j + 1] + g[7] * a1[j + 2];
                    U3(i, j) = -f;
                }
            } else {
#pragma unroll
                for (int k = 7; k >= 0; k--) {
                    const int j = NX - 1 - k;
                    if (j > c1) break;
                    const float f = g[0] * U3(i, j - 1) + g[1] * U3(i, j - 2) + g[2] * a2[k] + g[3] * a2[k + 1] +
                                    g[4] * a2[k + 2] + g[5] * a1[k] + g[6] * a1[k + 1] + g[7] * a1[k + 2];
                    U3(i, j) = -f;
                }
            }
        }
        __syncthreads();
    }
}

template <bool ADJ>
__device__ __forceinline__ void fd_step2_body(const StepArgs &A, const ShotDev &S, const int lb) {
    constexpr int NT = kThreads2;
    constexpr int NA = (R0X * R0Z + NT - 1) / NT;   // U_n cells per thread
    constexpr int N1 = (R1X * R1Z + NT - 1) / NT;   // U_{n+1} cells per thread
    constexpr int N2 = B2X * B2Z / NT;              // U_{n+2} cells per thread
    __shared__ float s_a[R0X * R0Z];             // U_n (v^2 U_n for the adjoint-style stencil)
    __shared__ float s_b[R1X * R1Z];             // U_{n+1}
    __shared__ float s_c[ADJ ? R1X * R1Z : 1];   // v^2 U_{n+1} (adjoint-style stencil)
    __shared__ float s_d[B2X * B2Z];             // U_{n+2}
    __shared__ int s_inj[R1X * R1Z];             // per region cell: (first contribution << 8) | count
    __shared__ float s_trv[2][kMaxTileTraces2];  // both steps' trace samples
    __shared__ float s_red[B2X + B2X / 4];   // FP16 snapshot scales: column maxima, block scales

    const int its1 = A.dir > 0 ? A.n + 1 : S.nt - A.n;
    int it1 = its1;
    if (A.nbase) {
        const int b = *gp(A.nbase);
        it1 = A.dir > 0 ? its1 + b : its1 - b;
    }
    const int its2 = A.dir > 0 ? its1 + 1 : its1 - 1, it2 = A.dir > 0 ? it1 + 1 : it1 - 1;
    const int NZ = S.NZ, NX = S.NX, LDZ = S.LDZ, zoff = S.zoff;
    const int t = lb - S.tile0;
    const int z0 = (t % S.ntz) * B2Z - S.zsh, x0 = (t / S.ntz) * B2X - S.xsh;
    const GLB float *__restrict__ un = gp(S.u[A.rot]);
    const GLB float *__restrict__ un1 = gp(S.u[lev(A.rot, -1)]);
    const GLB float *__restrict__ un2 = gp(S.u[lev(A.rot, -2)]);
    const GLB float *__restrict__ un3 = gp(S.u[lev(A.rot, -3)]);
    GLB float *__restrict__ w1 = gp(S.u[lev(A.rot, 1)]);
    GLB float *__restrict__ w2 = gp(S.u[lev(A.rot, 2)]);
    const GLB float *__restrict__ v2p = gp(S.v2);
    const int tid = threadIdx.x;
    const float *c = A.c;
    auto adr = [&](int iz, int ix) -> unsigned {
        const int izc = iz < 0 ? 0 : (iz >= NZ ? NZ - 1 : iz), ixc = ix < 0 ? 0 : (ix >= NX ? NX - 1 : ix);
        return (unsigned)ixc * (unsigned)LDZ + (unsigned)(zoff + izc);
    };

    // ---- every global operand of step n is requested up front: U_n (and v^2)
    // on rows z0-8..z0+71 x columns x0-8..x0+39, then U_{n-1..n-3} and v^2 of
    // this thread's U_{n+1} cells.  Clamped addresses: outside the array the
    // values only feed cells that are dropped.
    float ra[NA], rva[ADJ ? NA : 1];
#pragma unroll
    for (int k = 0; k < NA; k++) {
        const int l = tid + k * NT;
        const unsigned q = adr(z0 - 8 + l % R0Z, x0 - 8 + l / R0Z);
        ra[k] = l < R0X * R0Z ? ldw(un, q) : 0.f;
        if (ADJ) rva[k] = l < R0X * R0Z ? ldw(v2p, q) : 0.f;
    }
    float p1[N1], p0[N1], pm[N1], pv[N1], pu[ADJ ? N1 : 1];
#pragma unroll
    for (int k = 0; k < N1; k++) {
        const int l = tid + k * NT;
        const unsigned q = adr(z0 - 4 + l % R1Z, x0 - 4 + l / R1Z);
        const bool ok = l < R1X * R1Z;
        p1[k] = ok ? ldw(un1, q) : 0.f;
        p0[k] = ok ? ldw(un2, q) : 0.f;
        pm[k] = ok ? ldw(un3, q) : 0.f;
        pv[k] = ok ? ldw(v2p, q) : 0.f;
        if (ADJ) pu[k] = ok ? ldw(un, q) : 0.f;
    }
#pragma unroll
    for (int k = 0; k < NA; k++) {
        const int l = tid + k * NT;
        if (l < R0X * R0Z) s_a[l] = ADJ ? rva[k] * ra[k] : ra[k];
    }
    // ---- injection plan of the region and both steps' trace samples
    const GLB int *ti = gp(S.inj2_tile);
    const int has_inj = A.inject && ti[t + 1] > ti[t];
    if (has_inj) {
        for (int l = tid; l < R1X * R1Z; l += NT) s_inj[l] = 0;
        const GLB int *ttr = gp(S.inj2_ttr);
        const GLB float *trs = gp(S.inj_traces);
        const int j0 = gp(S.inj2_tptr)[t], j1 = gp(S.inj2_tptr)[t + 1];
        for (int j = j0 + tid; j < j1; j += NT) {
            s_trv[0][j - j0] = trs[(size_t)ttr[j] * S.nt + (it1 - 1)];
            s_trv[1][j - j0] = trs[(size_t)ttr[j] * S.nt + (it2 - 1)];
        }
        __syncthreads();
        const GLB int *cptr = gp(S.inj2_cptr), *cell = gp(S.inj2_cell);
        for (int k = ti[t] + tid; k < ti[t + 1]; k += NT) {
            const int a = cptr[k], b = cptr[k + 1];
            s_inj[cell[k]] = (a << 8) | (b - a);
        }
    }
    __syncthreads();
    const GLB float4v *co = (const GLB float4v *)gp(S.inj2_coef);

    // ---- step n: extraction, snapshot, imaging on u2 = U_n
    auto un_own = [&](int lz, int lx) -> float {
        return ADJ ? ldw(un, adr(z0 + lz, x0 + lx)) : s_a[(lz + 8) + R0Z * (lx + 8)];
    };
    if (A.extract && S.ext_tile && gp(S.ext_tile)[t + 1] > gp(S.ext_tile)[t])
        extract_lds(S, t, it1, s_a, z0 - 8, x0 - 8, R0Z, z0, x0);
    if ((its1 % A.snapi) == 0 && (A.store_snap || (A.imaging != IMG_NONE && A.imaging != IMG_BORN)))
        snap_image_tile(S, A, t, it1, z0, x0, s_red, un_own);

    // ---- U_{n+1} on rows z0-4..z0+67, columns x0-4..x0+35
#pragma unroll
    for (int k = 0; k < N1; k++) {
        const int l = tid + k * NT;
        if (l >= R1X * R1Z) break;
        const int lz = l % R1Z, lx = l / R1Z, iz = z0 - 4 + lz, ix = x0 - 4 + lx;
        float v = 0.f;
        if (iz >= 4 && iz <= NZ - 5 && ix >= 4 && ix <= NX - 5) {
            const float *s = s_a + (lz + 4) + R0Z * (lx + 4);
            const float tz = c[0] * (s[0]) + c[1] * (s[1] + s[-1]) + c[2] * (s[2] + s[-2]) + c[3] * (s[3] + s[-3]) +
                             c[4] * (s[4] + s[-4]);
            const float tx = c[5] * (s[0]) + c[6] * (s[R0Z] + s[-R0Z]) + c[7] * (s[2 * R0Z] + s[-2 * R0Z]) +
                             c[8] * (s[3 * R0Z] + s[-3 * R0Z]) + c[9] * (s[4 * R0Z] + s[-4 * R0Z]);
            float val = ADJ ? (tx + tz) : pv[k] * (tx + tz);
            if (has_inj) {
                const int e = s_inj[l];
                if (e & 0xff) val = inject_cell(val, co, e >> 8, e & 0xff, s_trv[0]);
            }
            const float u2c = ADJ ? pu[k] : s[0];
            v = (20.f * u2c - 6.f * p1[k] - 4.f * p0[k] + pm[k] + A.c12dt2 * val) * A.div11;
        }
        s_b[l] = v;
    }
    // operands of step n+1 on the tile, requested before the band work
    float q1[ADJ ? N2 : 1], q0[N2], qm[N2], qv[N2];
#pragma unroll
    for (int k = 0; k < N2; k++) {
        const int l = tid + k * NT;
        const unsigned q = adr(z0 + l % B2Z, x0 + l / B2Z);
        if (ADJ) q1[k] = ldw(un, q);
        q0[k] = ldw(un1, q);
        qm[k] = ldw(un2, q);
        qv[k] = ldw(v2p, q);
    }
    __syncthreads();
    higdon_region(
        S, z0 - 4, z0 + R1Z - 5, x0 - 4, x0 + R1X - 5,
        [&](int i, int j) -> float & { return s_b[(i - (z0 - 4)) + R1Z * (j - (x0 - 4))]; },
        [&](int i, int j) -> float {
            return ADJ ? ldw(un, adr(i, j)) : s_a[(i - (z0 - 8)) + R0Z * (j - (x0 - 8))];
        },
        [&](int i, int j) -> float { return ldw(un1, adr(i, j)); });

    // ---- store U_{n+1} on the tile (corners are never written)
    for (int l = tid; l < B2X * B2Z; l += NT) {
        const int lz = l % B2Z, lx = l / B2Z, iz = z0 + lz, ix = x0 + lx;
        if (iz < 0 || iz >= NZ || ix < 0 || ix >= NX) continue;
        if (!(iz >= 4 && iz <= NZ - 5) && !(ix >= 4 && ix <= NX - 5)) continue;
        stw(w1, (unsigned)ix * (unsigned)LDZ + (unsigned)(zoff + iz), s_b[(lz + 4) + R1Z * (lx + 4)]);
    }
    if (ADJ) {
#pragma unroll
        for (int k = 0; k < N1; k++) {
            const int l = tid + k * NT;
            if (l < R1X * R1Z) s_c[l] = pv[k] * s_b[l];
        }
    }

    // ---- step n+1: extraction, snapshot, imaging on u2 = U_{n+1}
    auto un1_own = [&](int lz, int lx) -> float { return s_b[(lz + 4) + R1Z * (lx + 4)]; };
    if (A.extract && S.ext_tile && gp(S.ext_tile)[t + 1] > gp(S.ext_tile)[t])
        extract_lds(S, t, it2, s_b, z0 - 4, x0 - 4, R1Z, z0, x0);
    if ((its2 % A.snapi) == 0 && (A.store_snap || (A.imaging != IMG_NONE && A.imaging != IMG_BORN)))
        snap_image_tile(S, A, t, it2, z0, x0, s_red, un1_own);
    __syncthreads();

    // ---- U_{n+2} on the tile
    const float *src = ADJ ? s_c : s_b;
#pragma unroll
    for (int k = 0; k < N2; k++) {
        const int l = tid + k * NT;
        const int lz = l % B2Z, lx = l / B2Z, iz = z0 + lz, ix = x0 + lx;
        float v = 0.f;
        if (iz >= 4 && iz <= NZ - 5 && ix >= 4 && ix <= NX - 5) {
            const int b0 = (lz + 4) + R1Z * (lx + 4);
            const float *s = src + b0;
            const float tz = c[0] * (s[0]) + c[1] * (s[1] + s[-1]) + c[2] * (s[2] + s[-2]) + c[3] * (s[3] + s[-3]) +
                             c[4] * (s[4] + s[-4]);
            const float tx = c[5] * (s[0]) + c[6] * (s[R1Z] + s[-R1Z]) + c[7] * (s[2 * R1Z] + s[-2 * R1Z]) +
                             c[8] * (s[3 * R1Z] + s[-3 * R1Z]) + c[9] * (s[4 * R1Z] + s[-4 * R1Z]);
            float val = ADJ ? (tx + tz) : qv[k] * (tx + tz);
            if (has_inj) {
                const int e = s_inj[b0];
                if (e & 0xff) val = inject_cell(val, co, e >> 8, e & 0xff, s_trv[1]);
            }
            const float u2c = s_b[b0];
            const float r1 = ADJ ? q1[k] : s_a[(lz + 8) + R0Z * (lx + 8)];
            v = (20.f * u2c - 6.f * r1 - 4.f * q0[k] + qm[k] + A.c12dt2 * val) * A.div11;
        }
        s_d[l] = v;
    }
    __syncthreads();
    const bool edge = z0 <= 0 || z0 + B2Z >= NZ || x0 <= 0 || x0 + B2X >= NX;
    if (edge)
        higdon_region(
            S, z0, z0 + B2Z - 1, x0, x0 + B2X - 1,
            [&](int i, int j) -> float & { return s_d[(i - z0) + B2Z * (j - x0)]; },
            [&](int i, int j) -> float { return s_b[(i - (z0 - 4)) + R1Z * (j - (x0 - 4))]; },
            [&](int i, int j) -> float {
                return ADJ ? ldw(un, adr(i, j)) : s_a[(i - (z0 - 8)) + R0Z * (j - (x0 - 8))];
            });
    for (int l = tid; l < B2X * B2Z; l += NT) {
        const int lz = l % B2Z, lx = l / B2Z, iz = z0 + lz, ix = x0 + lx;
        if (iz < 0 || iz >= NZ || ix < 0 || ix >= NX) continue;
        if (!(iz >= 4 && iz <= NZ - 5) && !(ix >= 4 && ix <= NX - 5)) continue;
        stw(w2, (unsigned)ix * (unsigned)LDZ + (unsigned)(zoff + iz), s_d[l]);
    }
}

template <bool ADJ, bool SINGLE>
__global__ void __launch_bounds__(kThreads2) fd_step2_kernel(StepArgs A) {
    const int lb = logical_block(A);
    if (lb >= A.nblocks) return;
    if constexpr (SINGLE) {
        fd_step2_body<ADJ>(A, A.s0, lb);
    } else {
        fd_step2_body<ADJ>(A, A.shots[gp(A.blk_shot)[lb]], lb);
    }
}

}  // namespace sfd
