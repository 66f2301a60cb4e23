// fd_step2.h — time blocking: two steps of the time loop in one launch.
//
// A workgroup owns a 56 x 32 tile.  Its ten wavefronts each own a 64-row x
// 4-column strip of the tile's U_{n+1} region — the tile and a 4-cell ring
// (64 x 40) — with lane = region row, so one wavefront spans the tile's rows
// and both 4-row rings.  Step n runs as in the one-step kernel (x-window of
// U_n in registers, z-stencil through a wavefront-private LDS column, the
// strip's U_{n-1}, U_{n-2}, U_{n-3}, v^2 in registers, injection, 5-level
// update) and leaves U_{n+1} of the region in LDS; after one workgroup
// barrier (and the Higdon bands of edge regions) the eight wavefronts of the
// tile's own strips compute U_{n+2} from that LDS copy, reusing the U_n,
// U_{n-1}, U_{n-2} and v^2 they already hold for the same cells.  U_n ..
// U_{n-3} and v^2 cross HBM once for two steps; the ring costs 25% more
// stencil work in step n only.  Per cell the arithmetic and its order are the
// one-step kernel's, so the results are bitwise equal to two one-step
// launches.  Extraction, snapshots and imaging of both steps act on the
// tile's own cells (FP16 scale blocks of 56 x 4 cells, the tiling's).  Born
// passes, free surfaces and ghosts run the one-step kernel.
//
// Injection uses a per-tile plan over the 64 x 40 region (the tile's own
// cells and the ring cells of its neighbours' plans), built on the host.
#pragma once
#include <hip/hip_runtime.h>

namespace sfd {

constexpr int B2Z = 56, B2X = 32;               // own tile (rows, columns)
constexpr int R1Z = B2Z + 8, R1X = B2X + 8;     // U_{n+1} region (ring 4): 64 x 40
constexpr int kW2 = R1X / 4;                    // wavefronts: one per 4-column strip of the region
constexpr int kThreads2 = 64 * kW2;
constexpr int kMaxTileTraces2 = 512;            // distinct traces one region may inject
constexpr int kCellAhead = 2;                   // contributions per cell loaded before use
static_assert(R1Z == 64, "a wavefront's lanes span the region's rows");

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

// Higdon bands (Boundary.f90:119-153) of the lines crossing rows [r0, r1] x
// columns [c0, c1]: z-bands, then x-bands, each line a sequential recursion in
// the reference's order, computed only on the part of the line inside the
// region (its cells read the line's next two cells, which lie inside too).
// U3 / U2 / U1 give the new level, u2 and u1.  Ends with a workgroup barrier.
template <typename F3, typename F2, typename F1>
__device__ __forceinline__ void higdon_region(const ShotDev &S, int r0, int r1, int c0, int c1, F3 U3, F2 U2,
                                              F1 U1) {
    const int NZ = S.NZ, NX = S.NX, tid = threadIdx.x;
    const int nc = c1 - c0 + 1, nr = r1 - r0 + 1;
    const bool top = r0 <= 7 && !S.no_top, bot = r1 >= NZ - 8, left = c0 <= 7, right = c1 >= NX - 8;
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
            if (i < S.hrow0 || i > NZ - 5) continue;
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
    }
    __syncthreads();
}

// Snapshot and imaging of one step on a wavefront's strip (own cells only):
// u holds u2 of the strip's PX columns in this lane's row.  The FP16 scale of
// the strip's 56 x 4 block is a wavefront reduction.
template <int PX>
__device__ __forceinline__ void snap_image_strip(const ShotDev &S, const StepArgs &A, int it, int iz, int xs,
                                                 int tzi, bool own_row, const float *u) {
    const int lane = threadIdx.x & 63;
    const bool interior_z = own_row && iz >= S.iz0 && iz < S.iz0 + S.nz;
    float inv16 = 0.f;
    if (A.store_snap && A.snap_fp16) {
        float m = 0.f;
#pragma unroll
        for (int p = 0; p < PX; p++) {
            const int ix = xs + p;
            if (interior_z && ix >= S.i0 && ix < S.i0 + S.nxw) m = fmaxf(m, fabsf(u[p]));
        }
        for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
        if (!(m > 1e-30f)) m = 0.f;
        if (lane == 0)
            gp(S.snap_scale)[(size_t)(it / A.snapi - 1) * S.nsb + tzi + (size_t)S.ntz * ((xs + S.xsh) >> 2)] = m;
        inv16 = m > 0.f ? 1.0f / m : 0.f;
    }
    const bool img = A.imaging == IMG_RTM || A.imaging == IMG_FWI || A.imaging == IMG_LSRTM;
    const size_t npix = (size_t)S.nz * S.nxw;
#pragma unroll
    for (int p = 0; p < PX; p++) {
        const int ix = xs + p;
        if (!(interior_z && ix >= S.i0 && ix < S.i0 + S.nxw)) continue;
        const int sb = snap_block(S, iz, ix);
        float Sv = 0.f, img0 = 0.f, ill0 = 0.f, tp = 0.f, vel = 0.f;
        if (img) {
            const size_t wq = (size_t)(ix - S.i0) * S.nz + (iz - S.iz0);
            Sv = load_snap(S, A.snap_fp16, it / A.snapi, npix, wq, sb, true);
            img0 = gp(S.image)[wq];
            ill0 = gp(S.illum)[wq];
            tp = gp(S.taper)[ix - S.i0];
            if (A.imaging == IMG_FWI || A.imaging == IMG_LSRTM) vel = gp(S.vel)[(size_t)ix * S.LDZ + S.zoff + iz];
        }
        if (A.imaging == IMG_FWI)
            snap_image_point<true>(S, A, it, iz, ix, u[p], vel, sb, inv16, Sv, img0, ill0, tp);
        else
            snap_image_point<false>(S, A, it, iz, ix, u[p], vel, sb, inv16, Sv, img0, ill0, tp);
    }
}

template <bool ADJ>
__device__ __forceinline__ void fd_step2_body(const StepArgs &A, const ShotDev &S, const int lb) {
    constexpr int NT = kThreads2, PX = 4, NWIN = PX + 8;
    __shared__ float s_col[kW2][R1Z + 8];       // per-wavefront z column of U_n: the strip's rows + 4 + 4 halo
    __shared__ float s_b[R1X * R1Z];            // U_{n+1} on the region, column-major
    __shared__ float s_c[ADJ ? R1X * R1Z : 1];  // v^2 U_{n+1} (adjoint-style stencil)
    __shared__ float s_d[B2X * B2Z];            // U_{n+2} of edge tiles, for the Higdon bands
    __shared__ int s_inj[R1X * R1Z];            // per region cell: (first contribution << 8) | count
    __shared__ float s_trv[2][kMaxTileTraces2]; // both steps' trace samples

    const int its1 = A.dir > 0 ? A.n + 1 : S.nt - A.n;
    int it1 = its1;
    if (A.nbase) {
        const int b = *gp(A.nbase);
        it1 = A.dir > 0 ? its1 + b : its1 - b;
    }
    const int its2 = A.dir > 0 ? its1 + 1 : its1 - 1, it2 = A.dir > 0 ? it1 + 1 : it1 - 1;
    const int NZ = S.NZ, NX = S.NX, LDZ = S.LDZ, zoff = S.zoff;
    const int t = lb - S.tile0;
    const int tzi = t % S.ntz, txi = t / S.ntz;
    const int z0 = tzi * B2Z - S.zsh, x0 = txi * B2X - S.xsh;
    const GLB float *__restrict__ un = gp(S.u[A.rot]);
    const GLB float *__restrict__ un1 = gp(S.u[lev(A.rot, -1)]);
    const GLB float *__restrict__ un2 = gp(S.u[lev(A.rot, -2)]);
    const GLB float *__restrict__ un3 = gp(S.u[lev(A.rot, -3)]);
    GLB float *__restrict__ w1 = gp(S.u[lev(A.rot, 1)]);
    GLB float *__restrict__ w2 = gp(S.u[lev(A.rot, 2)]);
    const GLB float *__restrict__ v2p = gp(S.v2);
    const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const float *c = A.c;
    // lane = region row (rows z0-4 .. z0+59), wavefront wv = region columns 4wv .. 4wv+3
    const int iz = z0 - 4 + lane;
    const int sx = 4 * wv;                 // first region column of the strip
    const int xs = x0 - 4 + sx;            // its grid column
    const bool zarr = iz >= 0 && iz < NZ;
    const bool zin = iz >= 4 && iz <= NZ - 5;
    const bool own_row = lane >= 4 && lane < 4 + B2Z;
    const bool own_strip = wv >= 1 && wv <= kW2 - 2;   // wavefront-uniform
    // halo rows of the z-stencil served by lanes 0..7: 4 above, 4 below the strip
    const int hz = lane < 4 ? z0 - 8 + lane : z0 + 56 + lane;

    // ---- receivers of step n (U_n in memory)
    if (A.extract && S.ext_tile && t >= S.ext_lo && t < S.ext_hi && gp(S.ext_tile)[t + 1] > gp(S.ext_tile)[t])
        extract_tile(S, t, it1, un, z0, x0);

    // ---- loads, all issued before any use; clamped addresses only feed
    // cells outside the FD region, whose results are dropped
    const int izc = iz < 0 ? 0 : (iz >= NZ ? NZ - 1 : iz);
    const int hzc = hz < 0 ? 0 : (hz >= NZ ? NZ - 1 : hz);
    const unsigned zoc = (unsigned)(zoff + izc), hzo = (unsigned)(zoff + hzc);
    float wu[NWIN], wv2[ADJ ? NWIN : 1], hu[PX], hv[ADJ ? PX : 1];
    float r1[PX], r0[PX], rm[PX], rv[PX];
#pragma unroll
    for (int k = 0; k < NWIN; k++) {
        int ix = xs - 4 + k;
        ix = ix < 0 ? 0 : (ix >= NX ? NX - 1 : ix);
        wu[k] = ldw(un, (unsigned)ix * (unsigned)LDZ + zoc);
        if (ADJ) wv2[k] = ldw(v2p, (unsigned)ix * (unsigned)LDZ + zoc);
    }
#pragma unroll
    for (int p = 0; p < PX; p++) {
        int ix = xs + p;
        ix = ix < 0 ? 0 : (ix >= NX ? NX - 1 : ix);
        const unsigned cb = (unsigned)ix * (unsigned)LDZ;
        hu[p] = ldw(un, cb + hzo);
        if (ADJ) hv[p] = ldw(v2p, cb + hzo);
        r1[p] = ldw(un1, cb + zoc);
        r0[p] = ldw(un2, cb + zoc);
        // U_{n-3} is read by this launch only: non-temporal
        rm[p] = __builtin_nontemporal_load((const GLB float *)((const GLB char *)un3 + ((cb + zoc) << 2)));
        rv[p] = ldw(v2p, cb + zoc);
    }

    // ---- injection plan of the region and both steps' trace samples
    const GLB int *ti = gp(S.inj2_tile);
    const bool has_inj = A.inject && ti[t + 1] > ti[t];   // workgroup-uniform
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
        __syncthreads();
    }
    const GLB float4v *co = (const GLB float4v *)gp(S.inj2_coef);

    // ---- snapshot / imaging of step n on U_n (own cells)
    if ((its1 % A.snapi) == 0 && (A.store_snap || (A.imaging != IMG_NONE && A.imaging != IMG_BORN)) && own_strip)
        snap_image_strip<PX>(S, A, it1, iz, xs, tzi, own_row, wu + 4);

    // Adjoint: the stencil reads v^2 u (FDderivatives.f90:877-887)
    if (ADJ) {
#pragma unroll
        for (int k = 0; k < NWIN; k++) wv2[k] = wv2[k] * wu[k];
#pragma unroll
        for (int p = 0; p < PX; p++) hv[p] = hv[p] * hu[p];
    }

    // ---- step n: U_{n+1} on the strip (every region row)
    float *col = s_col[wv];
    float u3v[PX];
#pragma unroll
    for (int p = 0; p < PX; p++) {
        col[lane + 4] = ADJ ? wv2[p + 4] : wu[p + 4];
        if (lane < 8) col[lane < 4 ? lane : R1Z + lane] = ADJ ? hv[p] : hu[p];
        __builtin_amdgcn_wave_barrier();
        const float tz = c[0] * (col[lane + 4]) + c[1] * (col[lane + 5] + col[lane + 3]) +
                         c[2] * (col[lane + 6] + col[lane + 2]) + c[3] * (col[lane + 7] + col[lane + 1]) +
                         c[4] * (col[lane + 8] + col[lane]);
        __builtin_amdgcn_wave_barrier();
        const int ix = xs + p;
        float v = 0.f;
        if (zin && ix >= 4 && ix <= NX - 5) {
#define WX(k) (ADJ ? wv2[p + 4 + (k)] : wu[p + 4 + (k)])
            const float tx = c[5] * (WX(0)) + c[6] * (WX(1) + WX(-1)) + c[7] * (WX(2) + WX(-2)) +
                             c[8] * (WX(3) + WX(-3)) + c[9] * (WX(4) + WX(-4));
#undef WX
            float val = ADJ ? (tx + tz) : rv[p] * (tx + tz);
            if (has_inj) {
                const int e = s_inj[lane + R1Z * (sx + p)];
                if (e & 0xff) val = inject_cell(val, co, e >> 8, e & 0xff, s_trv[0]);
            }
            v = (20.f * wu[p + 4] - 6.f * r1[p] - 4.f * r0[p] + rm[p] + A.c12dt2 * val) * A.div11;
        }
        u3v[p] = v;
        s_b[lane + R1Z * (sx + p)] = v;
    }
    __syncthreads();

    // ---- Higdon bands of step n on the region (edge regions)
    const bool edge1 = (z0 - 4 <= 7 && !S.no_top) || z0 + 59 >= NZ - 8 || x0 - 4 <= 7 || x0 + 35 >= NX - 8;
    auto adr = [&](int i, int j) -> unsigned {
        const int ic = i < 0 ? 0 : (i >= NZ ? NZ - 1 : i), jc = j < 0 ? 0 : (j >= NX ? NX - 1 : j);
        return (unsigned)jc * (unsigned)LDZ + (unsigned)(zoff + ic);
    };
    if (edge1) {
        higdon_region(
            S, z0 - 4, z0 + R1Z - 5, x0 - 4, x0 + R1X - 5,
            [&](int i, int j) -> float & { return s_b[(i - (z0 - 4)) + R1Z * (j - (x0 - 4))]; },
            [&](int i, int j) -> float { return ldw(un, adr(i, j)); },
            [&](int i, int j) -> float { return ldw(un1, adr(i, j)); });
#pragma unroll
        for (int p = 0; p < PX; p++) u3v[p] = s_b[lane + R1Z * (sx + p)];
    }
    if (ADJ) {
#pragma unroll
        for (int p = 0; p < PX; p++) s_c[lane + R1Z * (sx + p)] = rv[p] * u3v[p];
        __syncthreads();
    }

    // ---- store U_{n+1} on the tile's cells (corners are never written)
    if (own_strip && own_row && zarr) {
#pragma unroll
        for (int p = 0; p < PX; p++) {
            const int ix = xs + p;
            if (ix < 0 || ix >= NX) continue;
            if (!zin && !(ix >= 4 && ix <= NX - 5)) continue;
            stw(w1, (unsigned)ix * (unsigned)LDZ + (unsigned)(zoff + iz), u3v[p]);
        }
    }

    // ---- step n+1: receivers and snapshot / imaging on U_{n+1}
    if (A.extract && S.ext_tile && t >= S.ext_lo && t < S.ext_hi && gp(S.ext_tile)[t + 1] > gp(S.ext_tile)[t])
        extract_lds(S, t, it2, s_b, z0 - 4, x0 - 4, R1Z, z0, x0);
    if ((its2 % A.snapi) == 0 && (A.store_snap || (A.imaging != IMG_NONE && A.imaging != IMG_BORN)) && own_strip)
        snap_image_strip<PX>(S, A, it2, iz, xs, tzi, own_row, u3v);

    // ---- U_{n+2} on the tile (own strips, own rows), from the LDS copy
    float u4[PX];
#pragma unroll
    for (int p = 0; p < PX; p++) u4[p] = 0.f;
    if (own_strip) {
        const float *src = ADJ ? s_c : s_b;
        const int lc = lane < 4 ? 4 : (lane > 3 + B2Z ? 3 + B2Z : lane);   // other lanes read valid slots
#pragma unroll
        for (int p = 0; p < PX; p++) {
            const float *s = src + lc + R1Z * (sx + p);
            const float tz = c[0] * (s[0]) + c[1] * (s[1] + s[-1]) + c[2] * (s[2] + s[-2]) + c[3] * (s[3] + s[-3]) +
                             c[4] * (s[4] + s[-4]);
            const float tx = c[5] * (s[0]) + c[6] * (s[R1Z] + s[-R1Z]) + c[7] * (s[2 * R1Z] + s[-2 * R1Z]) +
                             c[8] * (s[3 * R1Z] + s[-3 * R1Z]) + c[9] * (s[4 * R1Z] + s[-4 * R1Z]);
            const int ix = xs + p;
            if (own_row && zin && ix >= 4 && ix <= NX - 5) {
                float val = ADJ ? (tx + tz) : rv[p] * (tx + tz);
                if (has_inj) {
                    const int e = s_inj[lane + R1Z * (sx + p)];
                    if (e & 0xff) val = inject_cell(val, co, e >> 8, e & 0xff, s_trv[1]);
                }
                u4[p] = (20.f * u3v[p] - 6.f * wu[p + 4] - 4.f * r1[p] + r0[p] + A.c12dt2 * val) * A.div11;
            }
        }
    }
    // ---- Higdon bands of step n+1 on the tile (edge tiles)
    const bool edge2 = (z0 <= 7 && !S.no_top) || z0 + B2Z - 1 >= NZ - 8 || x0 <= 7 || x0 + B2X - 1 >= NX - 8;
    if (edge2) {
        if (own_strip && own_row) {
#pragma unroll
            for (int p = 0; p < PX; p++) s_d[(lane - 4) + B2Z * (sx - 4 + p)] = u4[p];
        }
        __syncthreads();
        higdon_region(
            S, z0, z0 + B2Z - 1, x0, x0 + B2X - 1,
            [&](int i, int j) -> float & { return s_d[(i - z0) + B2Z * (j - x0)]; },
            [&](int i, int j) -> float { return s_b[(i - (z0 - 4)) + R1Z * (j - (x0 - 4))]; },
            [&](int i, int j) -> float { return ldw(un, adr(i, j)); });
        if (own_strip && own_row) {
#pragma unroll
            for (int p = 0; p < PX; p++) u4[p] = s_d[(lane - 4) + B2Z * (sx - 4 + p)];
        }
    }
    if (own_strip && own_row && zarr) {
#pragma unroll
        for (int p = 0; p < PX; p++) {
            const int ix = xs + p;
            if (ix < 0 || ix >= NX) continue;
            if (!zin && !(ix >= 4 && ix <= NX - 5)) continue;
            stw(w2, (unsigned)ix * (unsigned)LDZ + (unsigned)(zoff + iz), u4[p]);
        }
    }
}

template <bool ADJ, bool SINGLE>
__global__ void __launch_bounds__(kThreads2) __attribute__((amdgpu_waves_per_eu(8))) fd_step2_kernel(StepArgs A) {
    const int lb = logical_block(A);
    if (lb >= A.nblocks) return;
    if constexpr (SINGLE) {
        fd_step2_body<ADJ>(A, A.s0, lb);
    } else {
        fd_step2_body<ADJ>(A, A.shots[gp(A.blk_shot)[lb]], lb);
    }
}

}  // namespace sfd
