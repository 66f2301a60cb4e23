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
constexpr int kMaxTileTraces = 1024;   // distinct traces one tile may inject
constexpr int kMaxTileCoef = 2048;     // injection contributions of one tile staged in LDS
constexpr int kLevels = 6;             // wavefield level buffers per shot (two-step launches write two)

// buffer of level U_{n+d} given rot = n mod kLevels
__host__ __device__ __forceinline__ int lev(int rot, int d) { return (rot + d + 2 * kLevels) % kLevels; }

// One injection contribution: gain (v^2 or 1/(dz dx)), sinc weights, and the
// tile-local index of the trace it injects.
struct __attribute__((aligned(16))) InjCoef {
    float a, wz, wx;
    int tr;
};
typedef float float4v __attribute__((ext_vector_type(4)));

// Global-address-space views of device pointers.  Pointers read from the
// descriptor structs are generic to the compiler, which then emits FLAT
// accesses; those also count against the LDS wait counter and serialise the
// streaming loads with the LDS stencil reads.  Casting to address space 1
// gives global_load/global_store.
#define GLB __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ GLB T *gp(T *p) { return (GLB T *)p; }
template <class T>
__device__ __forceinline__ const GLB T *gp(const T *p) { return (const GLB T *)p; }
// Wavefield element at a 32-bit element offset (< 2^30 cells per array):
// a uniform base plus a 32-bit byte offset maps to the SGPR-base addressing
// form, one VGPR per address instead of a 64-bit pair.
__device__ __forceinline__ float ldw(const GLB float *base, unsigned e) {
    return *(const GLB float *)((const GLB char *)base + (e << 2));
}
__device__ __forceinline__ void stw(GLB float *base, unsigned e, float v) {
    *(GLB float *)((GLB char *)base + (e << 2)) = v;
}

// Imaging modes (Imaging.f90)
enum { IMG_NONE = 0, IMG_RTM = 1, IMG_FWI = 2, IMG_LSRTM = 3, IMG_BORN = 4, IMG_RHOVP = 5 };
// Extraction modes (Extraction.f90), -1 none
enum { EXT_SINC_V2 = 0, EXT_NEAREST = 1, EXT_SINC_LSRTM = 2, EXT_SINC_UNIT = 3 };

// Per-shot device descriptor (all pointers are device pointers).
struct ShotDev {
    float *u[kLevels];      // wavefield levels, LDZ*NX each: U_j lives in u[j mod kLevels]
    const float *v2;        // vel*vel, padded
    const float *vel;       // vel, padded (imaging / Born)
    const float *hz;        // Higdon z-band coefficients: 16 per column ix in [4, NX-5]
    const float *hx;        // Higdon x-band coefficients: 16 per row iz in [4, NZ-5]
    int NZ, NX, LDZ, zoff;  // padded dims, leading dim, storage offset of row 0
    int zsh, xsh;           // tile-origin shifts (tile k starts at k*T - sh)
    int tz;                 // tile rows (64, or 56 for the two-step kernel's tiling)
    int ntz, ntx, tile0;    // tile grid and first logical tile
    int nz, nxw, i0, iz0;   // window interior, its first padded column (4+ntaper) and row (4+ntop;
                            // ntop = 1-nmin1 FD rows lie above model row 1)
    int nt, ntsnap;
    // injection: per tile [inj_tile[t], inj_tile[t+1]) cells; per cell a
    // contribution range [inj_cptr[c], inj_cptr[c+1])
    const int *inj_tile;
    const int *inj_cell;    // local cell index lz + TZ*lx
    const int *inj_cptr;
    const InjCoef *inj_coef; // per contribution
    const int *inj_tptr;    // per tile [inj_tptr[t], inj_tptr[t+1]) traces it injects
    const int *inj_ttr;     // global trace index of each tile-local trace
    int inj_lo, inj_hi;     // every injecting tile t lies in [inj_lo, inj_hi)
    // per tile 8 ints {first contribution, end, first tile-local trace (into
    // inj_ttr), end, offset of the tile's cell map in inj_map, the tile's
    // first global trace when its traces are consecutive (else -1), 0, 0}; the
    // cell map holds, for every cell of an injecting tile (lz + TZ*lx), its
    // contribution range relative to the tile's first: start << 8 | count
    const int *inj_desc;
    const int *inj_map;
    // the same injections planned per 72x40 region (tile + 4-cell ring) for
    // the two-step kernel: region-local cells, per-region trace tables
    const int *inj2_tile, *inj2_cell, *inj2_cptr, *inj2_tptr, *inj2_ttr;
    const InjCoef *inj2_coef;
    const float *inj_traces;// source traces, (nt, ntraces) time fastest
    float *inj_tm;          // the same traces time-major, [(it-1) * inj_ntr + trace]
    int inj_ntr;
    // extraction: per tile [ext_tile[t], ext_tile[t+1]) receivers
    const int *ext_tile;
    const int *ext_rec;     // receiver index
    const int *ext_pos;     // local centre cell (lz | lx << 16)
    const float *ext_w;     // 16 floats per receiver: gain, wz[7], wx[7], pad
    float *traces;          // extracted traces (nt, ntr)
    int ext_kind;
    int ext_lo, ext_hi;     // every extracting tile t lies in [ext_lo, ext_hi)
    // snapshots (window interior, nz*nxw per snapshot)
    void *snap;             // float* or __half*
    float *snap_scale;      // FP16 mode: max|S| over the window cells of each 64 x 4 block of the tile
                            // grid (rows of tiles, 4-column groups), nsb per snapshot
    int nsb;                // FP16 scale blocks per snapshot: ntz * ntx * (tile width / 4)
    // imaging / Born
    float *image, *illum;   // nz*nxw window accumulators
    const float *born_m;    // nz*nxw reflectivity
    const float *taper;     // nxw
    float dt2snap;          // (dt*snapi)^2
    // free surface and ghosts (Boundary.f90:1609-1658, Injection.f90:447-450,
    // Extraction.f90:548-556): padded row M - r mirrors row r (M = 2 iz0 - 2,
    // Fortran row -r of row r); fs_row[ix] is the padded row of the surface
    // cell ielev_z of FD column ix, fs_w[8 ix ..] its sinc weights and an
    // off-grid flag (w[7]); null when surf_type is 0
    int mirror;             // M
    int ext_ghost;          // rec_type != 0: receivers read u(r) - u(M - r)
    int no_top;             // surf_type > 0: no top Higdon band (Boundary.f90:123)
    int hrow0;              // first padded row of the side Higdon bands (nz0, Boundary.f90:104-108)
    int snap_mask;          // a window column's surface lies below model row 1: snapshots zero above it
    const int *fs_row;
    const float *fs_w;
    // variable density (withRho): v2 holds v^2 rho; delp = 1/rho2 (rho at the
    // half cells), which the step kernel's staggered gradients read; sz is
    // the set-up's scratch (rho2 along z)
    const float *delp;
    float *sz;
    const float *rho;       // variable density: rho, padded (the v-rho imaging condition reads it)
    float *image2;          // IMG_RHOVP: the rho gradient's window accumulator (nz*nxw)
};

struct StepArgs {
    ShotDev s0;             // the shot itself when the batch has one shot (no table lookups)
    int single;
    const ShotDev *shots;
    const int *blk_shot;    // logical tile -> shot
    const int *shot_order;  // shots by decreasing record length (image_kernel's z index -> shot)
    int nblocks;            // logical tiles of this launch: those of the shots still running,
                            // a prefix of the layout (grid may be padded to a multiple of 8)
    int xcd_remap;          // map block ids so one XCD gets contiguous tiles
    int step_grid;          // blocks of the launch that step tiles (the rest image, mode 3)
    int img_gx, img_gy;     // mode 3: column groups and 256-row chunks of the largest window
    int nshots;             // mode 3: shots of the batch
    int n;                  // step counter from 0 (plus *nbase when nbase is set)
    const int *nbase;       // device-held step base of a replayed graph, or null
    int rot;                // n mod kLevels: u2 = U_n is u[rot]
    int dir;                // +1 forward time, -1 reverse time
    int snapi;
    int store_snap;         // 1: store u2 as the snapshot of this time
    int snap_fp16;          // snapshots are FP16 (u / scale)
    int snap_global;        // FP16: one scale per snapshot (max |u| over the window, computed by
                            // snap_max_kernel / snap_scale_kernel before the launch) instead of one per block
    float *snap_part;       // snap_global: per-shot partial maxima (kSnapParts per place of the batch)
    int imaging;            // IMG_*
    int extract;            // 1: run receiver extraction
    int inject;             // 1: run injection
    float c[10];            // c0z..c4z, c0x..c4x (FDtypes.f90:367-387)
    float c12dt2;           // 12*dt^2
    float div11;            // 1/11
    float rdz, rdx;         // 1/dz, 1/dx: variable density's first-order padding gradients
};

__device__ __forceinline__ size_t addr(const ShotDev &S, int iz, int ix) {
    return (size_t)ix * (size_t)S.LDZ + (size_t)(S.zoff + iz);
}

// Logical tile of this workgroup, XCD-aware: consecutive logical tiles (which
// share halos) run on the same XCD and hit its L2.
__device__ __forceinline__ int logical_block(const StepArgs &A) {
    int b = blockIdx.x;
    if (A.xcd_remap) {
        int per = A.step_grid >> 3;
        b = (b & 7) * per + (b >> 3);
    }
    return b;
}

// Snapshot idx (1-based) at window offset wq, decoded to FP32.  stream: a
// non-temporal load, for the imaging point's own snapshot, read once by RTM,
// LSRTM and Born, so it does not push the wavefield levels of the next steps
// out of the caches.  FWI imaging also reads the neighbouring snapshots
// idx -+ 1 with cached loads; a snapshot's first read there is as a
// neighbour, which keeps it for the two imaging steps that follow.
// FP16 scale block of padded cell (iz, ix): 64 rows of a tile row by 4
// columns, numbered down the tile rows first.
__device__ __forceinline__ int snap_block(const ShotDev &S, int iz, int ix) {
    const int r = iz + S.zsh;
    return (S.tz == 64 ? r >> 6 : r / 56) + S.ntz * ((ix + S.xsh) >> 2);
}
__device__ __forceinline__ float load_snap(const ShotDev &S, int fp16, int idx, size_t npix, size_t wq, int sb,
                                           bool stream) {
    const size_t q = (size_t)(idx - 1) * npix + wq;
    if (fp16) {
        const GLB unsigned short *p = gp((const unsigned short *)S.snap) + q;
        const unsigned short h = stream ? __builtin_nontemporal_load(p) : *p;
        return __half2float(__ushort_as_half(h)) * gp(S.snap_scale)[(size_t)(idx - 1) * S.nsb + sb];
    }
    const GLB float *p = gp((const float *)S.snap) + q;
    return stream ? __builtin_nontemporal_load(p) : *p;
}

// Receivers whose centre lies in this tile (ExtractData on u2,
// Extraction.f90:539-545 / :84-86); the 7x7 footprint is read from memory.
__device__ __forceinline__ void extract_tile(const ShotDev &S, int t, int it, const GLB float *u2, int z0, int x0) {
    const GLB int *et = gp(S.ext_tile);
    for (int k = et[t] + (int)threadIdx.x; k < et[t + 1]; k += (int)blockDim.x) {
        const int r = gp(S.ext_rec)[k];
        const int pos = gp(S.ext_pos)[k];
        const int iz = z0 + (pos & 0xffff), ix = x0 + (pos >> 16);
        GLB float *tr = gp(S.traces) + (size_t)r * S.nt + (it - 1);
        if (S.ext_kind == EXT_NEAREST) {
            if (S.ext_ghost) *tr = *tr + (u2[addr(S, iz, ix)] - u2[addr(S, S.mirror - iz, ix)]);
            else *tr = *tr + u2[addr(S, iz, ix)];
        } else {
            // trace += g*u*wz*wx, j outer, i inner; with a receiver ghost
            // trace += g*(u(r) - u(mirror of r))*wz*wx (Extraction.f90:548-556)
            const GLB float *w = gp(S.ext_w) + 16 * k;
            const float g = w[0];
            float acc = *tr;
            if (S.ext_ghost) {
                for (int j = -3; j <= 3; j++)
                    for (int i = -3; i <= 3; i++)
                        acc = acc + g * (u2[addr(S, iz + i, ix + j)] - u2[addr(S, S.mirror - iz - i, ix + j)]) *
                                        w[4 + i] * w[11 + j];
            } else {
                for (int j = -3; j <= 3; j++)
                    for (int i = -3; i <= 3; i++) acc = acc + g * u2[addr(S, iz + i, ix + j)] * w[4 + i] * w[11 + j];
            }
            *tr = acc;
        }
    }
}

// The imaging conditions at one point: RTM img + u S t (Imaging.f90:265),
// FWI img + u (S- - 2S + S+) / (dt_snap^2 v^3) t (:365), LSRTM
// img + 2 t u S / v (:738).
template <bool FWI>
__device__ __forceinline__ float image_value(int imaging, float img, float u2c, float Sv, float Sm, float Sp,
                                             float vel, float tp, float dt2snap) {
    if (FWI) {
        const float tmp = (Sm - 2.f * Sv + Sp) / (dt2snap * (vel * vel * vel));
        return img + u2c * tmp * tp;
    }
    if (imaging == IMG_RTM) return img + u2c * Sv * tp;
    return img + 2.f * tp * u2c * Sv / vel;
}

// Snapshot store and imaging condition at one interior point (on u2, before
// the stencil: Extraction.f90:595, Imaging.f90:265 / :365 / :738).
// Sv, img0, ill0, tp: the source snapshot, image, illumination and taper at
// this point, loaded ahead by the caller (imaging modes only).
// FWI: the FWI imaging condition (the launch images with IMG_FWI); otherwise
// RTM or LSRTM as A.imaging says (nothing for other values).
template <bool FWI>
__device__ __forceinline__ void snap_image_point(const ShotDev &S, const StepArgs &A, int it, int iz, int ix,
                                                 float u2c, float vel, int sb, float inv16, float Sv, float img0,
                                                 float ill0, float tp) {
    const int wz = iz - S.iz0, wx = ix - S.i0;
    const size_t wq = (size_t)wx * S.nz + wz;
    const size_t npix = (size_t)S.nz * S.nxw;
    const int idx = it / A.snapi;
    if (A.store_snap) {
        if (A.snap_fp16) {
            __builtin_nontemporal_store(__half_as_ushort(__float2half_rn(u2c * inv16)),
                                        gp((unsigned short *)S.snap) + (size_t)(idx - 1) * npix + wq);
        } else {
            __builtin_nontemporal_store(u2c, gp((float *)S.snap) + (size_t)(idx - 1) * npix + wq);
        }
    }
    if (FWI || A.imaging == IMG_RTM || A.imaging == IMG_LSRTM) {
        GLB float *imp = gp(S.image), *ilp = gp(S.illum);
        float Sm = 0.f, Sp = 0.f;
        if (FWI) {
            Sm = load_snap(S, A.snap_fp16, idx - 1 > 1 ? idx - 1 : 1, npix, wq, sb, false);
            Sp = load_snap(S, A.snap_fp16, idx + 1 < S.ntsnap ? idx + 1 : S.ntsnap, npix, wq, sb, false);
        }
        imp[wq] = image_value<FWI>(A.imaging, img0, u2c, Sv, Sm, Sp, vel, tp, S.dt2snap);
        ilp[wq] = ill0 + Sv * Sv * tp;
    }
}

// Imaging at one step of a reverse pass (Imaging.f90:228-275 RTM, :322-370
// FWI, :698-745 LSRTM) on the u2 that the step tiles of the same launch read
// and do not write: 256 threads take 256 rows (chunk gy) of COLS window
// columns (group gx) and issue all loads of their cells before any use.
// These blocks follow the step tiles in the launch (mode 3), so they start
// in the slots the first finished tiles free and the step tiles keep their
// occupancy.
// Columns per thread, measured on the 751x2301 RTM and the 96-shot batch:
// two for the imaging blocks of a one-shot launch, four for image_kernel.
constexpr int kImgColsSingle = 2, kImgColsBatch = 4;
template <bool FWI, int COLS>
__device__ __forceinline__ void image_cells(const StepArgs &A, const ShotDev &S, int gx, int gy, int lane256) {
    const int wx0 = gx * COLS, wz = gy * 256 + lane256;
    if (wx0 >= S.nxw || wz >= S.nz) return;
    const int its = A.dir > 0 ? A.n + 1 : S.nt - A.n;
    int it = its;
    if (A.nbase) {
        const int b = *gp(A.nbase);
        it = A.dir > 0 ? its + b : its - b;
    } else if (it < 1 || it > S.nt) {
        return;   // this shot has finished (ragged batch)
    }
    if (its % A.snapi != 0) return;
    const int idx = it / A.snapi, iz = wz + S.iz0;
    const size_t npix = (size_t)S.nz * S.nxw;
    const GLB float *u2 = gp(S.u[A.rot]);
    const bool vel_needed = FWI || A.imaging == IMG_LSRTM;
    float u[COLS], Sv[COLS], Sm[COLS], Sp[COLS], I0[COLS], L0[COLS], T[COLS], V[COLS];
#pragma unroll
    for (int k = 0; k < COLS; k++) {
        const int wx = wx0 + k < S.nxw ? wx0 + k : S.nxw - 1;   // clamped: loads stay in the window
        const int ix = wx + S.i0;
        const size_t q = (size_t)wx * S.nz + wz;
        const size_t a = (size_t)ix * S.LDZ + S.zoff + iz;
        const int t = snap_block(S, iz, ix);
        u[k] = u2[a];
        Sv[k] = load_snap(S, A.snap_fp16, idx, npix, q, t, true);
        if (FWI) {
            Sm[k] = load_snap(S, A.snap_fp16, idx - 1 > 1 ? idx - 1 : 1, npix, q, t, false);
            Sp[k] = load_snap(S, A.snap_fp16, idx + 1 < S.ntsnap ? idx + 1 : S.ntsnap, npix, q, t, false);
        } else {
            Sm[k] = Sp[k] = 0.f;
        }
        I0[k] = gp(S.image)[q];
        L0[k] = gp(S.illum)[q];
        T[k] = gp(S.taper)[wx];
        V[k] = vel_needed ? gp(S.vel)[a] : 0.f;
    }
#pragma unroll
    for (int k = 0; k < COLS; k++) {
        if (wx0 + k >= S.nxw) break;
        const size_t q = (size_t)(wx0 + k) * S.nz + wz;
        gp(S.image)[q] = image_value<FWI>(A.imaging, I0[k], u[k], Sv[k], Sm[k], Sp[k], V[k], T[k], S.dt2snap);
        gp(S.illum)[q] = L0[k] + Sv[k] * Sv[k] * T[k];
    }
}

// Batched launches image in a kernel of their own beside the lean step
// (grid: column groups, 256-row chunks, the shots still running in order of
// decreasing record length): folded into the step launch, the batched
// variant's shot-table reads spill registers.
template <bool FWI>
__global__ void __launch_bounds__(256) image_kernel(StepArgs A) {
    image_cells<FWI, kImgColsBatch>(A, A.shots[gp(A.shot_order)[blockIdx.z]], (int)blockIdx.x, (int)blockIdx.y,
                                    (int)threadIdx.x);
}

// Imaging_condition_AFWI_RHOVP_from_memory (Imaging.f90:424-512) in 2D at
// one step of a reverse pass with variable density: one thread per window
// cell (grid: window columns, 256-row chunks, the running shots in order of
// decreasing record length), launched before the free surface and the step
// change u2.  With ip = min(nz, i+1), jp = min(j+1, nxw), im = ip-1,
// jm = jp-1 (window cells, 1-based):
//   tmp  = dzi (u(ip,jm) - u(im,jm)) dzi (S(ip,jm) - S(im,jm))
//        + dxi (u(im,jp) - u(im,jm)) dxi (S(im,jp) - S(im,jm))  (+ the 2D y term, 0)
//   tmp1 = u(i,j) / (dt snapi)^2 (S- - 2 S + S+)
//   image  += 2 tmp1 t / (v^3 rho),  image2 += tmp t / rho^2 + tmp1 t / (rho^2 v^2),
//   illum  += S S t
// in the reference's operation order; rows above the free surface are not
// imaged (:455-458).
__global__ void __launch_bounds__(256) image_rhovp_kernel(StepArgs A, float dzi, float dxi) {
    const ShotDev &S = A.shots[gp(A.shot_order)[blockIdx.z]];
    const int wx = blockIdx.x, wz = blockIdx.y * 256 + threadIdx.x;
    if (wx >= S.nxw || wz >= S.nz) return;
    const int its = A.dir > 0 ? A.n + 1 : S.nt - A.n;
    int it = its;
    if (A.nbase) {
        const int b = *gp(A.nbase);
        it = A.dir > 0 ? its + b : its - b;
    } else if (it < 1 || it > S.nt) {
        return;   // this shot has finished (ragged batch)
    }
    if (its % A.snapi != 0) return;
    const int iz = wz + S.iz0, ix = wx + S.i0;
    if (S.fs_row && iz < gp(S.fs_row)[ix]) return;
    const int idx = it / A.snapi;
    const size_t npix = (size_t)S.nz * S.nxw;
    const int i = wz + 1, j = wx + 1;
    const int ip = i + 1 < S.nz ? i + 1 : S.nz, jp = j + 1 < S.nxw ? j + 1 : S.nxw;
    const int im = ip - 1, jm = jp - 1;
    const GLB float *u2 = gp(S.u[A.rot]);
    auto U = [&](int a, int b) { return u2[(size_t)(b - 1 + S.i0) * S.LDZ + S.zoff + (a - 1 + S.iz0)]; };
    auto W = [&](int k, int a, int b) {
        return load_snap(S, A.snap_fp16, k, npix, (size_t)(b - 1) * S.nz + (a - 1),
                         snap_block(S, a - 1 + S.iz0, b - 1 + S.i0), false);
    };
    const float uu = U(i, j);
    const float Sv = W(idx, i, j);
    const float Sm = W(idx - 1 > 1 ? idx - 1 : 1, i, j), Sp = W(idx + 1 < S.ntsnap ? idx + 1 : S.ntsnap, i, j);
    const float wmm = W(idx, im, jm);
    float tmp = dzi * (U(ip, jm) - U(im, jm)) * dzi * (W(idx, ip, jm) - wmm);
    tmp = tmp + dxi * (U(im, jp) - U(im, jm)) * dxi * (W(idx, im, jp) - wmm);
    tmp = tmp + 0.f;
    const float tmp1 = uu * (1.f / S.dt2snap) * (Sm - 2.f * Sv + Sp);
    const size_t a = (size_t)ix * S.LDZ + S.zoff + iz, q = (size_t)wx * S.nz + wz;
    const float v = gp(S.vel)[a], r = gp(S.rho)[a], t = gp(S.taper)[wx];
    GLB float *img = gp(S.image), *img2 = gp(S.image2), *ill = gp(S.illum);
    img[q] = img[q] + 2.f * tmp1 * t / (v * v * v * r);
    img2[q] = img2[q] + tmp * t / (r * r) + tmp1 * t / (r * r * (v * v));
    ill[q] = ill[q] + Sv * Sv * t;
}

// Boundary_set_free_surface_grid (Boundary.f90:1609-1658) on u2 of every
// running shot, after extraction, snapshot and imaging have read it and
// before the stencil does: one thread per FD column (grid: column groups,
// the running shots in order of decreasing record length).  Rows nmin1 ..
// ielev-1 are zeroed; an on-grid surface then odd-mirrors rows ielev ..
// ielev+2 into ielev-2 .. ielev-4; an off-grid one subtracts the mirror image
// of the sinc-interpolated column, in the reference's order.  Only rows above
// the surface change, and only from rows at and below it.
__global__ void __launch_bounds__(256) free_surface_kernel(StepArgs A) {
    const ShotDev &S = A.single ? A.s0 : A.shots[gp(A.shot_order)[blockIdx.y]];
    const int ix = 4 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (ix > S.NX - 5) return;
    const int its = A.dir > 0 ? A.n + 1 : S.nt - A.n;
    if (!A.nbase && (its < 1 || its > S.nt)) return;
    GLB float *u = gp(S.u[A.rot]) + (size_t)ix * S.LDZ + S.zoff;
    const int E = gp(S.fs_row)[ix];
    for (int r = 4; r < E; r++) u[r] = 0.f;
    const GLB float *w = gp(S.fs_w) + 8 * ix;
    if (w[7] == 0.f) {
#pragma unroll
        for (int i = 0; i <= 2; i++) u[E - 2 - i] = -u[E + i];
    } else {
        float wl[7], ui[7];
#pragma unroll
        for (int l = 0; l < 7; l++) wl[l] = w[l];
#pragma unroll
        for (int i = -3; i <= 3; i++) {
            float f = 0.f;
#pragma unroll
            for (int l = -3; l <= 3; l++) f = f + u[E + l + i] * wl[3 + l];
            ui[3 + i] = f;
        }
#pragma unroll
        for (int i = 2; i <= 4; i++)
#pragma unroll
            for (int l = -3; l <= i - 1; l++) u[E - i + l] = u[E - i + l] - ui[i + 1] * wl[3 + l];
    }
}


// Window + edge pad of the full-model density for one shot (as vel):
// rho into out.  Then rho_half_kernel (twice: along z into tmp, along x into
// delp) puts it on the half cells with the 9-point sinc at shift 0.5, the
// padded array's ends extended by their edge values (Interpolate.f90:44-102),
// and the x pass stores 1/rho2 (allocate_sxx2d_szz2d_delp2d, FDderivatives.f90:580).
__global__ void rho_pad_kernel(const float *__restrict__ big, int nz, const int *__restrict__ colsrc, int nxw,
                               int ntaper, int ntop, int NZ, int NX, int LDZ, int zoff, float *__restrict__ out) {
    const int iz = blockIdx.x * blockDim.x + threadIdx.x;
    const int ix = blockIdx.y;
    if (iz >= NZ || ix >= NX) return;
    int wz = iz - (4 + ntop), wx = ix - (4 + ntaper);
    wz = wz < 0 ? 0 : (wz >= nz ? nz - 1 : wz);
    wx = wx < 0 ? 0 : (wx >= nxw ? nxw - 1 : wx);
    out[(size_t)ix * LDZ + zoff + iz] = (float)0 + big[(size_t)colsrc[wx] * nz + wz];
}
struct Sinc9 { float w[9]; };
__global__ void rho_half_kernel(const float *__restrict__ in, float *__restrict__ out, int NZ, int NX, int LDZ,
                                int zoff, Sinc9 s, int along_x) {
    const int iz = blockIdx.x * blockDim.x + threadIdx.x;
    const int ix = blockIdx.y;
    if (iz >= NZ || ix >= NX) return;
    float f = 0.f;
#pragma unroll
    for (int l = -4; l <= 4; l++) {
        int i = along_x ? iz : iz + l, j = along_x ? ix + l : ix;
        i = i < 0 ? 0 : (i >= NZ ? NZ - 1 : i);
        j = j < 0 ? 0 : (j >= NX ? NX - 1 : j);
        f = f + s.w[4 + l] * in[(size_t)j * LDZ + zoff + i];
    }
    out[(size_t)ix * LDZ + zoff + iz] = along_x ? 1.f / f : f;
}
// v2 <- v2 * rho (the variable-density gain v^2 rho, FDderivatives.f90:651-652)
__global__ void v2_rho_kernel(float *__restrict__ v2, const float *__restrict__ rho, size_t n) {
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        v2[q] = v2[q] * rho[q];
}

// Receiver extraction on u2 before the free surface changes it (ExtractData
// precedes Boundary_set_free_surface_grid, Propagator.f90:176-198): launches
// with a free surface extract here, one block per logical tile of width TX,
// and the step kernel of the same step skips extraction.
__global__ void __launch_bounds__(256) extract_kernel(StepArgs A, int TX) {
    const int lb = blockIdx.x;
    if (lb >= A.nblocks) return;
    const ShotDev &S = A.single ? A.s0 : A.shots[gp(A.blk_shot)[lb]];
    const int t = lb - S.tile0;
    if (!S.ext_tile || t < S.ext_lo || t >= S.ext_hi) return;
    if (gp(S.ext_tile)[t + 1] <= gp(S.ext_tile)[t]) return;
    int it = A.dir > 0 ? A.n + 1 : S.nt - A.n;
    if (A.nbase) {
        const int b = *gp(A.nbase);
        it = A.dir > 0 ? it + b : it - b;
    } else if (it < 1 || it > S.nt) {
        return;
    }
    const int tzi = t % S.ntz, txi = t / S.ntz;
    extract_tile(S, t, it, gp((const float *)S.u[A.rot]), tzi * S.tz - S.zsh, txi * TX - S.xsh);
}

// Higdon absorbing boundary (Boundary.f90:119-153) on a tile that touches
// the grid edge.  s_b holds the tile's u3, s_u2 / s_u1 its u2 / u1, column
// stride TP = TZ + 1 (a thread walking a column and the lanes of a row then
// fall on distinct LDS banks).  z-bands first (one thread per column with ix
// in [nmin2, nmax2]), then x-bands (one thread per row with iz in [nmin1,
// nmax1]), each a sequential recursion outwards, exactly in the reference's
// order.  The coefficients of a thread's lines are loaded before the tile is
// staged (higdon_load_g) so their latency overlaps the staging barrier.
struct HigdonG {
    float z[8], x[8];
};

template <int TZ, int TX>
__device__ __forceinline__ void higdon_load_g(const ShotDev &S, int z0, int x0, bool top, bool bot, bool left,
                                              bool right, HigdonG &hg) {
    const int tid = threadIdx.x, NZ = S.NZ, NX = S.NX;
    if ((top || bot) && tid < 2 * TX) {
        const bool dotop = tid < TX;
        const int ix = x0 + tid % TX;
        if ((dotop ? top : bot) && ix >= 4 && ix <= NX - 5) {
            const GLB float *g = gp(S.hz) + 16 * (ix - 4) + (dotop ? 0 : 8);
#pragma unroll
            for (int k = 0; k < 8; k++) hg.z[k] = g[k];
        }
    }
    if ((left || right) && tid < 2 * TZ) {
        const bool doleft = tid < TZ;
        const int i = z0 + tid % TZ;
        if ((doleft ? left : right) && i >= S.hrow0 && i <= NZ - 5) {
            const GLB float *g = gp(S.hx) + 16 * (i - 4) + (doleft ? 0 : 8);
#pragma unroll
            for (int k = 0; k < 8; k++) hg.x[k] = g[k];
        }
    }
}

// One band line in registers.  Cells c_0 .. c_9 run from the inner side
// outwards (c_k at LDS index base + k * step): c_0, c_1 are FD cells whose u3
// is final, c_2 .. c_9 the band cells, each overwritten from the two cells
// inside it: u3(c_k) = -(g0 u3(c_k-1) + g1 u3(c_k-2) + g2 u2(c_k) + g3 u2(c_k-1)
// + g4 u2(c_k-2) + g5 u1(c_k) + g6 u1(c_k-1) + g7 u1(c_k-2)), summed left to right.
__device__ __forceinline__ void higdon_line(const float *g, float *s_b, const float *s_u2, const float *s_u1,
                                            int base, int step) {
    float a[10], c[10], b[10];
#pragma unroll
    for (int k = 0; k < 10; k++) {
        a[k] = s_u2[base + k * step];
        c[k] = s_u1[base + k * step];
    }
    b[0] = s_b[base];
    b[1] = s_b[base + step];
#pragma unroll
    for (int k = 2; k < 10; k++) {
        const float f = g[0] * b[k - 1] + g[1] * b[k - 2] + g[2] * a[k] + g[3] * a[k - 1] + g[4] * a[k - 2] +
                        g[5] * c[k] + g[6] * c[k - 1] + g[7] * c[k - 2];
        b[k] = -f;
    }
#pragma unroll
    for (int k = 2; k < 10; k++) s_b[base + k * step] = b[k];
}

template <int TZ, int TX>
__device__ __forceinline__ void higdon_tile(const ShotDev &S, const HigdonG &hg, int z0, int x0, bool top,
                                            bool bot, bool left, bool right, float *s_b, const float *s_u2,
                                            const float *s_u1) {
    constexpr int TP = TZ + 1;
    const int NZ = S.NZ, NX = S.NX, tid = threadIdx.x;
#define L(i, j) ((i) - z0 + TP * ((j) - x0))
    if (top || bot) {
        if (tid < 2 * TX) {
            const bool dotop = tid < TX;
            const int ix = x0 + tid % TX;
            if ((dotop ? top : bot) && ix >= 4 && ix <= NX - 5) {
                if (dotop) higdon_line(hg.z, s_b, s_u2, s_u1, L(9, ix), -1);           // rows 9, 8 | 7 .. 0
                else higdon_line(hg.z, s_b, s_u2, s_u1, L(NZ - 10, ix), 1);            // rows NZ-10, NZ-9 | NZ-8 ..
            }
        }
        __syncthreads();
    }
    if (left || right) {
        if (tid < 2 * TZ) {
            const bool doleft = tid < TZ;
            const int i = z0 + tid % TZ;
            if ((doleft ? left : right) && i >= S.hrow0 && i <= NZ - 5) {
                if (doleft) higdon_line(hg.x, s_b, s_u2, s_u1, L(i, 9), -TP);          // columns 9, 8 | 7 .. 0
                else higdon_line(hg.x, s_b, s_u2, s_u1, L(i, NX - 10), TP);            // columns NX-10, NX-9 | ..
            }
        }
    }
#undef L
    __syncthreads();
}

// The step kernel body.  A workgroup owns a 64 x 4*PX tile; each of its four
// wavefronts owns a 64 x PX strip: lane = row, and the lane keeps u2 of its
// row for the strip's columns plus 4 on each side in registers (the
// x-derivative).  The z-derivative goes through a wavefront-private LDS
// column (the 64 rows plus 4+4 halo rows loaded by lanes 0..7), so interior
// tiles need no workgroup barrier.  Edge tiles (Higdon) and tiles with
// injections synchronise the workgroup.
template <int PX>
struct StepShape {
    static constexpr int TZ = 64;
    static constexpr int TX = PX >= 8 ? 4 * PX : 32;   // >= 19 columns: Higdon bands stay in one tile
    static constexpr int NW = TX / PX;                   // wavefronts per workgroup
    static constexpr int THREADS = 64 * NW;
    // 64 x 4 strips: three 8-wave workgroups per CU need <= 80 VGPRs (6 waves
    // per SIMD), four need <= 64 (8 waves; the lean variant's LDS allows four)
    static constexpr int WAVES_PER_EU = PX == 4 ? 6 : 1;
    static constexpr int WAVES_LEAN = PX == 4 ? 8 : 1;
};

// MODE selects the code compiled in: 0 (lean) has no snapshot, imaging or
// Born code and is launched for steps where no shot of the batch is at a
// multiple of snapi (three steps in four); 1 adds the snapshot store (forward
// passes that keep snapshots); 2 adds the Born source term (Born's
// scattering pass), whose operands take registers the other modes do not
// spend; 3 is mode 0 plus imaging blocks after the step tiles (imaging
// steps of reverse passes of one shot, image_cells; batches launch mode 0
// and image_kernel).
template <int PX, bool ADJ, int MODE, int TZ, bool RHO>
__device__ __forceinline__ void fd_step_body(const StepArgs &A, const ShotDev &S, const int lb) {
    static_assert(!(RHO && (ADJ || MODE == 2)), "variable density: forward-style stencil, no Born term");
    constexpr bool IMG = MODE == 2;
    constexpr int NW = StepShape<PX>::NW, TX = StepShape<PX>::TX, NWIN = PX + 8;
    constexpr int NT = StepShape<PX>::THREADS;
    static_assert(NT >= 2 * TX && NT >= 2 * TZ, "one thread per Higdon line");
    static_assert(TZ == 64 || TZ == 56, "tiles of 64 or 56 rows (a wavefront's lanes)");
    __shared__ float s_col[NW][TZ + 8]; // per-wavefront z-stencil column
    __shared__ float s_ucol[NW][RHO ? TZ + 15 : 1];   // variable density: u2 of a column, rows z0-7 .. z0+TZ+7
    // Injection tables are read only inside the column loop and the edge
    // buffers are written only after it (behind a barrier on injection
    // tiles), so the two phases share one LDS block.  Keeping the block small
    // is what lets four 512-thread workgroups share a CU.
    __shared__ union {
        struct {
            float trv[kMaxTileTraces];      // this step's trace samples
            float4v coef[kMaxTileCoef];     // the tile's contributions
            int cmap[IMG ? TX * TZ : 1];    // MODE 2: each thread's cell-map entries, parked
        } inj;
        struct {
            float b[TX * (TZ + 1)];         // u3, column stride TZ + 1
            float u2[TX * (TZ + 1)];        // u2
            float u1[TX * (TZ + 1)];        // u1
        } edge;
    } s_p;
    float *const s_trv = s_p.inj.trv;
    float4v *const s_coef = s_p.inj.coef;

    // Reference time index.  `its` is the part fixed at launch; a replayed
    // graph adds a device-held base (a multiple of kLevels and of snapi, so `its`
    // carries the residues the buffer rotation and snapshot tests need).
    const int its = A.dir > 0 ? A.n + 1 : S.nt - A.n;
    if (!A.nbase && (its < 1 || its > S.nt)) return;   // this shot has finished (ragged batch)
    // The time index it = its (+ base) is read only where it is used
    // (extraction and injection tiles, snapshot and Born steps): the base is
    // a scalar memory read, and scalar reads return out of order, so reading
    // it up front would hold every tile's kernel-argument reads, and with
    // them its streaming loads, for one memory round trip.
    auto time_index = [&]() -> int {
        if (!A.nbase) return its;
        const int b = *gp(A.nbase);
        return A.dir > 0 ? its + b : its - b;
    };
    const int NZ = S.NZ, NX = S.NX, LDZ = S.LDZ, zoff = S.zoff;
    const int t = lb - S.tile0;
    const int tzi = t % S.ntz, txi = t / S.ntz;
    const int z0 = tzi * TZ - S.zsh, x0 = txi * TX - S.xsh;
    const GLB float *__restrict__ u2 = gp(S.u[A.rot]);
    const GLB float *__restrict__ u1 = gp(S.u[lev(A.rot, -1)]);
    const GLB float *__restrict__ u0 = gp(S.u[lev(A.rot, -2)]);
    const GLB float *__restrict__ um = gp(S.u[lev(A.rot, -3)]);   // u_{-1}
    GLB float *__restrict__ u3 = gp(S.u[lev(A.rot, 1)]);
    const GLB float *__restrict__ v2p = gp(S.v2);
    // wv is the same on every lane of a wavefront; readfirstlane lets the
    // compiler keep per-column addresses and taper reads in scalar registers
    const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int iz = z0 + lane;
    const int xs = x0 + wv * PX;                      // first column of this wavefront's strip
    const bool lane_ok = lane < TZ;                   // 56-row tiles leave lanes 56..63 idle
    const bool zarr = lane_ok && iz >= 0 && iz < NZ;
    const bool zin = lane_ok && iz >= 4 && iz <= NZ - 5;
    const size_t zo = (size_t)(zoff + iz);
    // halo row served by lanes 0..7: 4 above, 4 below the strip
    const int hz = lane < 4 ? z0 - 4 + lane : z0 + TZ - 4 + lane;
    const bool hok = lane < 8 && hz >= 0 && hz < NZ;

    // ---- receivers centred in this tile (reads u2 only; done before the
    // streaming loads so its registers are not live alongside them)
    if (A.extract && S.ext_tile && t >= S.ext_lo && t < S.ext_hi && gp(S.ext_tile)[t + 1] > gp(S.ext_tile)[t])
        extract_tile(S, t, time_index(), u2, z0, x0);

    // ---- injection tiles: descriptor, this step's trace samples, the lane's
    // cell-map entries and the first chunk of contributions (memory to LDS)
    // are requested before the streaming loads (a wave's loads return in
    // order, so these come back first); the samples reach LDS further down,
    // and the contributions are folded after the stencil pass.
    bool has_inj = A.inject && t >= S.inj_lo && t < S.inj_hi;
    int ic0 = 0, ic1 = 0, it0 = 0, it1 = 0, imap = 0;
    constexpr int kTrPerThread = (kMaxTileTraces + NT - 1) / NT;
    float trv_r[kTrPerThread];
    int me[PX];
    if (has_inj) {
        // tile-uniform: readfirstlane keeps them in scalar registers
        const GLB int *dd = gp(S.inj_desc) + 8 * t;
        ic0 = __builtin_amdgcn_readfirstlane(dd[0]);
        ic1 = __builtin_amdgcn_readfirstlane(dd[1]);
        it0 = __builtin_amdgcn_readfirstlane(dd[2]);
        it1 = __builtin_amdgcn_readfirstlane(dd[3]);
        imap = __builtin_amdgcn_readfirstlane(dd[4]);
        const int tbase = __builtin_amdgcn_readfirstlane(dd[5]);
        has_inj = ic1 > ic0;
        // injecting tiles carry the staging and the contribution loop on top
        // of a plain tile's work, and a launch ends with its slowest tile:
        // their waves issue first (measured: reverse-time launch -0.2 us on
        // a 2301-receiver line; the same for the Higdon edge tiles lost 1 us)
        if (has_inj) __builtin_amdgcn_s_setprio(2);
        if (has_inj) {
            const GLB float *tm = gp((const float *)S.inj_tm) + (size_t)(time_index() - 1) * S.inj_ntr;
            const GLB int *ttr = gp(S.inj_ttr);
#pragma unroll
            for (int r = 0; r < kTrPerThread; r++) {
                const int j = it0 + tid + r * NT;
                trv_r[r] = j < it1 ? tm[tbase >= 0 ? tbase + (j - it0) : ttr[j]] : 0.f;
            }
            const GLB int *mp = gp(S.inj_map) + imap;
#pragma unroll
            for (int p = 0; p < PX; p++) me[p] = lane_ok ? mp[lane + TZ * (wv * PX + p)] : 0;
            // The tile's contributions reach LDS in chunks of kMaxTileCoef
            // (one chunk except on the densest receiver tiles).  The first
            // chunk goes straight from memory into LDS (no registers), so
            // its loads can be issued here, ahead of the streaming loads:
            // each lane writes 16 B at the wave's base + lane * 16, which is
            // entry j for a wave whose base is entry j - lane.
            const GLB float4v *co = (const GLB float4v *)gp(S.inj_coef) + ic0;
            const int c1 = ic1 - ic0 < kMaxTileCoef ? ic1 - ic0 : kMaxTileCoef;
            for (int j = tid; j < c1; j += NT)
                __builtin_amdgcn_global_load_lds((const void *)(co + j),
                                                 (__attribute__((address_space(3))) void *)(s_coef + (j - lane)), 16, 0,
                                                 0);
        }
    }

    // ---- loads: u2 (and v^2 for the adjoint) window, halos, own u1/u0/u_-1/v^2.
    // Addresses are clamped into the array and the loads are unconditional,
    // so they issue back to back with no wait in between.  A clamped value
    // only ever feeds cells outside the FD region, whose results are dropped
    // (an FD cell's +-4 neighbours always lie inside the padded array).
    // (32-bit element offsets: one shot's wavefield is far below 2^32 cells,
    // and a uniform base plus a 32-bit offset needs one VGPR per address)
    const int izc = iz < 0 ? 0 : (iz >= NZ ? NZ - 1 : iz);
    const int hzc = hz < 0 ? 0 : (hz >= NZ ? NZ - 1 : hz);
    const unsigned zoc = (unsigned)(zoff + izc), hzo = (unsigned)(zoff + hzc);
    // Variable density: the stencil's operands are the staggered gradients
    // of u2 times delp = 1/rho2 (FDderivatives.f90:588-640), formed here from
    // u2 and delp (8th order on the FD region, first order times the boundary
    // row's or column's delp on the three padding rows and columns, zero on the
    // outermost row and column): sx along the lane's
    // row over the x-window (u2 of columns xs-7 .. xs+PX+7 in registers), sz
    // down each own column through an LDS column of u2 (below).
    constexpr int NU = RHO ? PX + 15 : 1;
    float wu[NWIN], wv2[ADJ ? NWIN : 1], hu[PX], hv[ADJ ? PX : 1], zc[RHO ? PX : 1], uo_r[RHO ? PX : 1];
    float ux[NU], dxw[RHO ? NWIN : 1], uh[RHO ? PX : 1], dzo[RHO ? PX : 1], dzh[RHO ? PX : 1];
    float r1[PX], r0[PX], rm[PX], rv[PX];
    if constexpr (RHO) {
        const GLB float *__restrict__ dp = gp(S.delp);
        // delp's row for a gradient at row r: the boundary row's on the padding rows
        auto drow = [&](int r) {
            const int q = r <= 3 ? 4 : (r >= NZ - 4 ? NZ - 5 : r);
            return q < 0 ? 0 : (q >= NZ ? NZ - 1 : q);
        };
        const unsigned dzoc = (unsigned)(zoff + drow(iz)), dhzo = (unsigned)(zoff + drow(hz));
        const int hr = lane < 7 ? z0 - 7 + lane : z0 + TZ + lane - 7;   // lanes 0..14: rows around the tile
        const unsigned hro = (unsigned)(zoff + (hr < 0 ? 0 : (hr >= NZ ? NZ - 1 : hr)));
#pragma unroll
        for (int m = 0; m < NU; m++) {
            int ix = xs - 7 + m;
            ix = ix < 0 ? 0 : (ix >= NX ? NX - 1 : ix);
            ux[m] = ldw(u2, (unsigned)ix * (unsigned)LDZ + zoc);
        }
#pragma unroll
        for (int k = 0; k < NWIN; k++) {
            // the padding columns take the boundary column's delp
            int jd = xs - 4 + k;
            jd = jd <= 3 ? 4 : (jd >= NX - 4 ? NX - 5 : jd);
            jd = jd < 0 ? 0 : (jd >= NX ? NX - 1 : jd);
            dxw[k] = ldw(dp, (unsigned)jd * (unsigned)LDZ + zoc);
        }
#pragma unroll
        for (int p = 0; p < PX; p++) {
            int ix = xs + p;
            ix = ix < 0 ? 0 : (ix >= NX ? NX - 1 : ix);
            const unsigned cb = (unsigned)ix * (unsigned)LDZ;
            uh[p] = ldw(u2, cb + hro);
            dzo[p] = ldw(dp, cb + dzoc);
            dzh[p] = ldw(dp, cb + dhzo);
            uo_r[p] = ux[p + 7];
        }
    } else {
#pragma unroll
        for (int k = 0; k < NWIN; k++) {
            int ix = xs - 4 + k;
            ix = ix < 0 ? 0 : (ix >= NX ? NX - 1 : ix);
            wu[k] = ldw(u2, (unsigned)ix * (unsigned)LDZ + zoc);
            if (ADJ) wv2[k] = ldw(v2p, (unsigned)ix * (unsigned)LDZ + zoc);
        }
    }
#pragma unroll
    for (int p = 0; p < PX; p++) {
        int ix = xs + p;
        ix = ix < 0 ? 0 : (ix >= NX ? NX - 1 : ix);
        const unsigned cb = (unsigned)ix * (unsigned)LDZ;
        if (!RHO) hu[p] = ldw(u2, cb + hzo);
        if (ADJ) hv[p] = ldw(v2p, cb + hzo);
        r1[p] = ldw(u1, cb + zoc);
        r0[p] = ldw(u0, cb + zoc);
        // u_-1 is read by this step only (two steps on, its buffer takes a
        // new level): a non-temporal load leaves the caches to the others
        rm[p] = __builtin_nontemporal_load((const GLB float *)((const GLB char *)um + ((cb + zoc) << 2)));
        if (!ADJ) rv[p] = ldw(v2p, cb + zoc);
    }
    (void)hok;
    // Adjoint: the stencil reads v^2 u2 (FDderivatives.f90:877-887); form the
    // products once so the outer window cells keep one register, not two.
    if (ADJ) {
#pragma unroll
        for (int k = 0; k < NWIN; k++) wv2[k] = wv2[k] * wu[k];
#pragma unroll
        for (int p = 0; p < PX; p++) hv[p] = hv[p] * hu[p];
    }

    const int it = (MODE == 1 || MODE == 2) ? time_index() : its;   // snapshot and Born steps
    // u2 of the lane's own cells
    float uo[PX];
#pragma unroll
    for (int p = 0; p < PX; p++) uo[p] = RHO ? uo_r[p] : wu[p + 4];

    // ---- snapshot steps of a forward pass that keeps snapshots (mode 1;
    // reverse-pass imaging runs in the imaging blocks of mode 3)
    const bool snap_step = MODE == 1 && (its % A.snapi) == 0 && A.store_snap;
    const bool interior_z = lane_ok && iz >= S.iz0 && iz < S.iz0 + S.nz;

    // ---- injection tiles (workgroup-uniform branch): the trace samples and
    // the tile's contributions staged in LDS, so a contribution costs one
    // packed load and one LDS read.  Modes 0 and 1 keep the cell-map entries
    // in registers; the imaging mode, whose registers hold the imaging
    // operands, parks them in thread-private LDS slots (read back by the same
    // thread: no barrier).
    if (has_inj) {
        if constexpr (IMG) {
#pragma unroll
            for (int p = 0; p < PX; p++)
                if (lane_ok) s_p.inj.cmap[lane + TZ * (wv * PX + p)] = me[p];
        }
#pragma unroll
        for (int r = 0; r < kTrPerThread; r++) {
            const int j = tid + r * NT;
            if (it0 + j < it1) s_trv[j] = trv_r[r];
        }
        // the staged entries are folded after the stencil pass, which so
        // runs while the memory-to-LDS loads are still arriving
    }

    // FP16 snapshots carry one scale per 64 x 4 block (a wavefront's strip
    // holds PX / 4 of them): max|u2| over the block's window cells at this
    // step, so values are stored as u2/max in [-1, 1].  A wavefront reduction
    // finds it: no LDS and no workgroup barrier on snapshot steps.
    constexpr int NB = PX / 4;
    float inv16[NB];
#pragma unroll
    for (int q = 0; q < NB; q++) inv16[q] = 0.f;
    if (snap_step && A.store_snap && A.snap_fp16) {
        const size_t sbase = (size_t)(it / A.snapi - 1) * S.nsb + tzi + (size_t)S.ntz * ((xs + S.xsh) >> 2);
#pragma unroll
        for (int q = 0; q < NB; q++) {
            float m;
            if (A.snap_global) {
                // one scale per snapshot, already in every block's slot
                m = gp(S.snap_scale)[sbase + (size_t)S.ntz * q];
            } else {
                m = 0.f;
#pragma unroll
                for (int p = 4 * q; p < 4 * q + 4; p++) {
                    const int ix = xs + p;
                    if (interior_z && ix >= S.i0 && ix < S.i0 + S.nxw &&
                        !(S.snap_mask && iz < gp(S.fs_row)[ix]))
                        m = fmaxf(m, fabsf(uo[p]));
                }
                for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
                // blocks whose largest value is below 1e-30 store zeros (1/m would overflow FP32)
                if (!(m > 1e-30f)) m = 0.f;
                if (lane == 0) gp(S.snap_scale)[sbase + (size_t)S.ntz * q] = m;
            }
            inv16[q] = m > 0.f ? 1.0f / m : 0.f;
        }
    }
    const float *c = A.c;
    if constexpr (RHO) {
        // sx on the lane's row over the x-window: 8th order on the FD columns,
        // first order on the three padding columns each side, zero beyond
        // (rows other than FD rows feed no result)
#pragma unroll
        for (int k = 0; k < NWIN; k++) {
            const int j = xs - 4 + k;
#define UXW(d) ux[k + 3 + (d)]
            float gx = 0.f;
            if (j >= 4 && j <= NX - 5)
                gx = (c[6] * (UXW(1) - UXW(0)) + c[7] * (UXW(2) - UXW(-1)) + c[8] * (UXW(3) - UXW(-2)) +
                      c[9] * (UXW(4) - UXW(-3))) * dxw[k];
            else if (j >= 1 && j <= NX - 2)
                gx = A.rdx * (UXW(1) - UXW(0)) * dxw[k];
#undef UXW
            wu[k] = gx;
        }
    }
    float *col = s_col[wv];
    const int cl = lane_ok ? lane : TZ - 1;   // idle lanes read a valid slot

    // Three passes over the strip's PX cells of a lane: stencil (and the
    // snapshot / imaging point), then the injections of all PX cells
    // interleaved (each cell's contributions still in its own order, but the
    // PX chains of LDS reads overlap), then the Born term and the time update.
    float val[PX];
#pragma unroll
    for (int p = 0; p < PX; p++) {
        const int ix = xs + p;
        if constexpr (RHO) {
            // sz of this column on the lane's row and (lanes 0..7) its halo row,
            // from u2 rows z0-7 .. z0+TZ+7 staged in the wavefront's LDS column
            float *uc = s_ucol[wv];
            if (lane_ok) uc[lane + 7] = ux[p + 7];
            if (lane < 15) uc[lane < 7 ? lane : TZ + lane] = uh[p];
            __builtin_amdgcn_wave_barrier();
            auto szat = [&](int i, int sl, float d) -> float {
                if (i >= 4 && i <= NZ - 5)
                    return (c[1] * (uc[sl + 1] - uc[sl]) + c[2] * (uc[sl + 2] - uc[sl - 1]) +
                            c[3] * (uc[sl + 3] - uc[sl - 2]) + c[4] * (uc[sl + 4] - uc[sl - 3])) * d;
                if (i >= 1 && i <= NZ - 2) return A.rdz * (uc[sl + 1] - uc[sl]) * d;
                return 0.f;
            };
            zc[p] = lane_ok ? szat(iz, lane + 7, dzo[p]) : 0.f;
            if (lane < 8) hu[p] = szat(hz, lane < 4 ? lane + 3 : TZ + lane + 3, dzh[p]);
            __builtin_amdgcn_wave_barrier();
        }
        const float u2c = RHO ? zc[p] : wu[p + 4];
        // z-stencil column through wavefront-private LDS (all 64 lanes take part)
        if (lane_ok) col[lane + 4] = ADJ ? wv2[p + 4] : u2c;
        if (lane < 8) col[lane < 4 ? lane : TZ + lane] = ADJ ? hv[p] : hu[p];
        __builtin_amdgcn_wave_barrier();
        float tz;
        if (RHO)   // staggered divergence of sz (FDderivatives.f90:641-644)
            tz = c[1] * (col[cl + 4] - col[cl + 3]) + c[2] * (col[cl + 5] - col[cl + 2]) +
                 c[3] * (col[cl + 6] - col[cl + 1]) + c[4] * (col[cl + 7] - col[cl]);
        else
            tz = c[0] * (col[cl + 4]) + c[1] * (col[cl + 5] + col[cl + 3]) + c[2] * (col[cl + 6] + col[cl + 2]) +
                 c[3] * (col[cl + 7] + col[cl + 1]) + c[4] * (col[cl + 8] + col[cl]);
        __builtin_amdgcn_wave_barrier();
        val[p] = 0.f;
        if (!(zin && ix >= 4 && ix <= NX - 5)) continue;
        // FD_scheme: FDderivatives.f90:728-738 (forward) / :877-887 (adjoint) / :645-652 (density)
#define WX(k) (ADJ ? wv2[p + 4 + (k)] : wu[p + 4 + (k)])
        float tx;
        if (RHO)
            tx = c[6] * (WX(0) - WX(-1)) + c[7] * (WX(1) - WX(-2)) + c[8] * (WX(2) - WX(-3)) + c[9] * (WX(3) - WX(-4));
        else
            tx = c[5] * (WX(0)) + c[6] * (WX(1) + WX(-1)) + c[7] * (WX(2) + WX(-2)) + c[8] * (WX(3) + WX(-3)) +
                 c[9] * (WX(4) + WX(-4));
#undef WX
        val[p] = ADJ ? (tx + tz) : rv[p] * (tx + tz);
    }
    // Snapshot store (Extraction.f90:595, u2 of this step on the window).
    // FP32 snapshots are stored non-temporal (read back only in the reverse
    // pass, they should not push the wavefield levels out of the caches);
    // FP16 ones with plain stores, whose 128-byte wavefront segments the L2
    // merges (on the 96-shot batch: forward pass 1731 against 1749 ms
    // non-temporal, reverse 2198 against 2185, the call 0.2% faster).
    if (snap_step && interior_z) {
        const size_t sbase = (size_t)(it / A.snapi - 1) * S.nz * S.nxw + (iz - S.iz0);
#pragma unroll
        for (int p = 0; p < PX; p++) {
            const int ix = xs + p;
            if (ix < S.i0 || ix >= S.i0 + S.nxw) continue;
            const size_t q = sbase + (size_t)(ix - S.i0) * S.nz;
            // with a free surface the snapshot is zero above it (Extraction.f90:580-592)
            const float su = S.snap_mask && iz < gp(S.fs_row)[ix] ? 0.f : uo[p];
            if (A.snap_fp16)
                gp((unsigned short *)S.snap)[q] = __half_as_ushort(__float2half_rn(su * inv16[p / 4]));
            else
                __builtin_nontemporal_store(su, gp((float *)S.snap) + q);
        }
    }

    // Injection: per-cell contributions in trace order (Injection.f90:434-440).
    // Only FD cells carry contributions (the plan rejects footprints outside).
    if (has_inj) {
        int kb[PX], kn[PX];
#pragma unroll
        for (int p = 0; p < PX; p++) {
            const int e = IMG ? (lane_ok ? s_p.inj.cmap[lane + TZ * (wv * PX + p)] : 0) : me[p];
            kb[p] = e >> 8;
            kn[p] = e & 0xff;
        }
        // Chunk by chunk (the first one staged and folded above), each
        // cell's contributions in its own order; the PX cells' LDS reads are
        // unconditional so they issue together (a cell past its count reads
        // entry 0 and its sum is not taken).
        const int nco = ic1 - ic0;
        {
            // the memory-to-LDS loads of the first chunk count in vmcnt
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __syncthreads();
            // Fold this step's sample into each staged contribution: slot 0
            // becomes the whole term ((gain * sample) * sinc_z) * sinc_x,
            // the reference's product in its order, so a cell reads one LDS
            // value per contribution, with no dependent gather, and adds it.
            // Each thread folds the entries its own lane staged.
            const int c1 = nco < kMaxTileCoef ? nco : kMaxTileCoef;
            for (int j = tid; j < c1; j += NT) {
                const float4v w = s_coef[j];
                s_coef[j][0] = w[0] * s_trv[__float_as_int(w[3])] * w[1] * w[2];
            }
        }
        if (nco <= kMaxTileCoef) {
            int km = 0;
#pragma unroll
            for (int p = 0; p < PX; p++) km = kn[p] > km ? kn[p] : km;
            __syncthreads();
#pragma unroll 1
            for (int j = 0; j < km; j += 2) {
                float c[PX], c2[PX];
#pragma unroll
                for (int p = 0; p < PX; p++) {
                    c[p] = s_coef[j < kn[p] ? kb[p] + j : 0][0];
                    c2[p] = s_coef[j + 1 < kn[p] ? kb[p] + j + 1 : 0][0];
                }
#pragma unroll
                for (int p = 0; p < PX; p++) {
                    val[p] = j < kn[p] ? val[p] + c[p] : val[p];
                    val[p] = j + 1 < kn[p] ? val[p] + c2[p] : val[p];
                }
            }
        } else {
            // more contributions than one chunk (dense receivers): chunk by
            // chunk, one cell at a time, in each cell's order
            const GLB float4v *co = (const GLB float4v *)gp(S.inj_coef) + ic0;
            for (int cb = 0; cb < nco; cb += kMaxTileCoef) {
                const int ce = nco - cb < kMaxTileCoef ? nco : cb + kMaxTileCoef;
                if (cb > 0) {
                    __syncthreads();   // every thread is done with the previous chunk
                    for (int j = cb + tid; j < ce; j += NT) {
                        float4v w = co[j];
                        w[0] = w[0] * s_trv[__float_as_int(w[3])] * w[1] * w[2];
                        s_coef[j - cb] = w;
                    }
                }
                __syncthreads();
#pragma unroll
                for (int p = 0; p < PX; p++) {
                    const int g0 = kb[p] > cb ? kb[p] : cb, g1 = kb[p] + kn[p] < ce ? kb[p] + kn[p] : ce;
                    for (int g = g0; g < g1; g++) val[p] = val[p] + s_coef[g - cb][0];
                }
            }
        }
    }
    float u3v[PX];
#pragma unroll
    for (int p = 0; p < PX; p++) {
        const int ix = xs + p;
        u3v[p] = 0.f;
        if (!(zin && ix >= 4 && ix <= NX - 5)) continue;
        float v = val[p];
        // Born secondary source: Imaging.f90:779-783, forward time only
        if (IMG && A.imaging == IMG_BORN && (its % A.snapi) == 0 && interior_z && ix >= S.i0 && ix < S.i0 + S.nxw) {
            const int wz = iz - S.iz0, wx = ix - S.i0;
            const size_t wq = (size_t)wx * S.nz + wz;
            const float Sv = load_snap(S, A.snap_fp16, it / A.snapi, (size_t)S.nz * S.nxw, wq, snap_block(S, iz, ix), true);
            v = v + 2.f * gp(S.taper)[wx] * gp(S.born_m)[wq] * Sv / gp(S.vel)[(size_t)ix * LDZ + zo];
        }
        // TimeDer: FDderivatives.f90:974-976
        u3v[p] = (20.f * uo[p] - 6.f * r1[p] - 4.f * r0[p] + rm[p] + A.c12dt2 * v) * A.div11;
    }

    const bool top = z0 <= 0 && !S.no_top, bot = z0 + TZ >= NZ, left = x0 <= 0, right = x0 + TX >= NX;
    if (top || bot || left || right) {
        HigdonG hg;
        higdon_load_g<TZ, TX>(S, z0, x0, top, bot, left, right, hg);
        if (has_inj) __syncthreads();   // injection tables read before the edge buffers overwrite them
        float *s_b = s_p.edge.b;
        constexpr int TP = TZ + 1;
#pragma unroll
        for (int p = 0; p < PX; p++) {
            const int l = lane + TP * (wv * PX + p);
            if (!lane_ok) continue;
            s_b[l] = u3v[p];
            s_p.edge.u2[l] = uo[p];
            s_p.edge.u1[l] = r1[p];
        }
        __syncthreads();
        higdon_tile<TZ, TX>(S, hg, z0, x0, top, bot, left, right, s_b, s_p.edge.u2, s_p.edge.u1);
#pragma unroll
        for (int p = 0; p < PX; p++) u3v[p] = s_b[cl + TP * (wv * PX + p)];
    }

    // ---- write u3 on every non-corner cell of the strip (FD region from the
    // time update, edge bands from Higdon).
    if (zarr) {
#pragma unroll
        for (int p = 0; p < PX; p++) {
            const int ix = xs + p;
            if (ix < 0 || ix >= NX) continue;
            const bool xband = ix >= 4 && ix <= NX - 5;
            if (!zin && !xband) continue;   // corner: never written by the reference
            stw(u3, (unsigned)ix * (unsigned)LDZ + (unsigned)(zoff + iz), u3v[p]);
        }
    }
}

// Kernel entry: SINGLE batches read the shot descriptor from the kernel
// arguments (scalar loads, no dependent table lookups); batched launches
// find their shot through the logical-tile table.
template <int PX, bool ADJ, bool SINGLE, int MODE, int TZ, bool RHO>
__global__ void __launch_bounds__(StepShape<PX>::THREADS)
__attribute__((amdgpu_waves_per_eu(RHO ? 4 : (MODE == 2 ? StepShape<PX>::WAVES_PER_EU : StepShape<PX>::WAVES_LEAN)))) fd_step_kernel(StepArgs A) {
    // mode 3: the lean step plus imaging blocks after the step tiles
    constexpr int BODY = MODE == 3 ? 0 : MODE;
    if constexpr (MODE == 3) {
        if ((int)blockIdx.x >= A.step_grid) {
            constexpr int NV = StepShape<PX>::THREADS / 256;   // 256-thread imaging groups per block
            const int vb = ((int)blockIdx.x - A.step_grid) * NV + (int)threadIdx.x / 256;
            const int gx = vb % A.img_gx, r = vb / A.img_gx, gy = r % A.img_gy, s = r / A.img_gy;
            if (s >= (SINGLE ? 1 : A.nshots)) return;
            const ShotDev &S = SINGLE ? A.s0 : A.shots[s];
            if (A.imaging == IMG_FWI)
                image_cells<true, kImgColsSingle>(A, S, gx, gy, (int)threadIdx.x & 255);
            else
                image_cells<false, kImgColsSingle>(A, S, gx, gy, (int)threadIdx.x & 255);
            return;
        }
    }
    const int lb = logical_block(A);
    if (lb >= A.nblocks) return;
    if constexpr (SINGLE) {
        fd_step_body<PX, ADJ, BODY, TZ, RHO>(A, A.s0, lb);
    } else {
        fd_step_body<PX, ADJ, BODY, TZ, RHO>(A, A.shots[gp(A.blk_shot)[lb]], lb);
    }
}

// Plain streaming copy used to calibrate hardware byte counters.
// The streaming ceiling of a 5-read / 1-write step at a given size: one point
// per thread, FP32 loads of five arrays and one store, no stencil (the step
// kernel's algorithmic traffic with nothing else).
__global__ void __launch_bounds__(256) stream5_calib_kernel(const float *__restrict__ a, const float *__restrict__ b,
                                                            const float *__restrict__ c, const float *__restrict__ d,
                                                            const float *__restrict__ e, float *__restrict__ o,
                                                            size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o[i] = (20.f * a[i] - 6.f * b[i] - 4.f * c[i] + d[i]) + e[i];
}

__global__ void copy_calib_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n4) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Edge-replicating window + pad of the full model for one shot
// (mod_window_pad ExtractPadModel.f90:509-530 + model_pad ReadSouVelRho.f90:545-596).
// colsrc[j-1] = model column (0-based) feeding window column j.
__global__ void window_pad_kernel(const float *__restrict__ big, int nz, const int *__restrict__ colsrc,
                                  int nxw, int ntaper, int ntop, int NZ, int NX, int LDZ, int zoff,
                                  float *__restrict__ vel, float *__restrict__ v2) {
    const int iz = blockIdx.x * blockDim.x + threadIdx.x;
    const int ix = blockIdx.y;
    if (iz >= NZ || ix >= NX) return;
    int wz = iz - (4 + ntop), wx = ix - (4 + ntaper);
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

// One FP16 scale per snapshot (SFD_SNAP_FP16G): max |u2| over a shot's window
// interior at its snapshot steps, in two passes without atomics.  Pass 1:
// grid (kSnapParts, running shots); block b of place z strides over the
// window columns and writes its maximum to snap_part[z * kSnapParts + b].
// Pass 2: one block per place reduces its partials and writes the scale
// (0 below 1e-30, as the per-block form) into every block slot of the
// snapshot, so the decoders read it as they read a per-block scale.
constexpr int kSnapParts = 64;
__device__ __forceinline__ int snap_step_of(const StepArgs &A, const ShotDev &S, int *idx) {
    const int its = A.dir > 0 ? A.n + 1 : S.nt - A.n;
    int it = its;
    if (A.nbase) {
        const int b = *gp(A.nbase);
        it = A.dir > 0 ? its + b : its - b;
    }
    if (its % A.snapi != 0 || it < 1 || it > S.nt) return 0;
    *idx = it / A.snapi;
    return 1;
}
__global__ void __launch_bounds__(256) snap_max_kernel(StepArgs A) {
    const int z = blockIdx.y;
    const ShotDev &S = A.single ? A.s0 : A.shots[gp(A.shot_order)[z]];
    int idx;
    float m = 0.f;
    if (snap_step_of(A, S, &idx)) {
        const GLB float *u2 = gp(S.u[A.rot]);
        for (int wx = blockIdx.x; wx < S.nxw; wx += gridDim.x) {
            const size_t c = (size_t)(wx + S.i0) * S.LDZ + S.zoff + S.iz0;
            const int w0 = S.snap_mask ? gp(S.fs_row)[wx + S.i0] - S.iz0 : 0;   // zero above a free surface
            for (int wz = threadIdx.x; wz < S.nz; wz += blockDim.x)
                if (wz >= w0) m = fmaxf(m, fabsf(u2[c + wz]));
        }
    }
    __shared__ float red[256];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + k]);
        __syncthreads();
    }
    if (threadIdx.x == 0) gp(A.snap_part)[(size_t)z * kSnapParts + blockIdx.x] = red[0];
}
__global__ void __launch_bounds__(256) snap_scale_kernel(StepArgs A) {
    const int z = blockIdx.x;
    const ShotDev &S = A.single ? A.s0 : A.shots[gp(A.shot_order)[z]];
    int idx;
    if (!snap_step_of(A, S, &idx)) return;
    float m = 0.f;
    for (int k = 0; k < kSnapParts; k++) m = fmaxf(m, gp(A.snap_part)[(size_t)z * kSnapParts + k]);
    if (!(m > 1e-30f)) m = 0.f;
    for (int b = threadIdx.x; b < S.nsb; b += blockDim.x) gp(S.snap_scale)[(size_t)(idx - 1) * S.nsb + b] = m;
}

// Zero n bytes at p: a grid-stride fill with 16-byte vector stores over the
// aligned middle and byte stores for the unaligned head and tail, used where
// a memset node must not be recorded.
__global__ void fill_zero_kernel(char *__restrict__ p, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x, t0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t head = (16 - ((uintptr_t)p & 15)) & 15;
    if (head > n) head = n;
    const size_t n16 = (n - head) / 16, tail = head + 16 * n16;
    for (size_t i = t0; i < head; i += stride) p[i] = 0;
    float4 *q = (float4 *)(p + head);
    for (size_t i = t0; i < n16; i += stride) q[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    for (size_t i = tail + t0; i < n; i += stride) p[i] = 0;
}

// Step base of replayed graphs (one thread; plain vector stores).
__global__ void set_base_kernel(int *nbase, int v) { *nbase = v; }

// src (ntr, nt) time fastest -> dst (nt, ntr) trace fastest, through 32 x 32
// LDS tiles (both sides coalesced).  Grid (ceil(nt/32), ceil(ntr/32)), 32 x 8.
__global__ void traces_tm_kernel(const float *__restrict__ src, float *__restrict__ dst, int nt, int ntr) {
    __shared__ float tile[32][33];
    const int t0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int r = r0 + k, it = t0 + threadIdx.x;
        tile[k][threadIdx.x] = (r < ntr && it < nt) ? src[(size_t)r * nt + it] : 0.f;
    }
    __syncthreads();
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int it = t0 + k, r = r0 + threadIdx.x;
        if (it < nt && r < ntr) dst[(size_t)it * ntr + r] = tile[threadIdx.x][k];
    }
}
__global__ void advance_base_kernel(int *nbase, int k) { *nbase = *nbase + k; }


// mod_copy_image (ExtractPadModel.f90:662-695): nearest-index scatter-add of
// every shot's window image into the full model.  One thread per model cell;
// for model column ix, inv_shot / inv_col [inv_ptr[ix], inv_ptr[ix+1]) list
// the (shot, window column) pairs that map to it in shot order, so the sum
// runs in shot order and starts from what `out` holds (calls for successive
// shot waves add in the same order as one call for all of them).
struct CopyImg {
    const float *img;       // nz*nxw
    const int *colmap;      // window column -> model column or -1
    int nxw;
};
__global__ void copy_image_kernel(const CopyImg *__restrict__ shots, int nz, int nx, const int *__restrict__ inv_ptr,
                                  const int *__restrict__ inv_shot, const int *__restrict__ inv_col,
                                  float *__restrict__ out) {
    const int iz = blockIdx.x * blockDim.x + threadIdx.x;
    const int ix = blockIdx.y;
    if (iz >= nz || ix >= nx) return;
    float acc = out[(size_t)ix * nz + iz];
    for (int k = inv_ptr[ix]; k < inv_ptr[ix + 1]; k++)
        acc = acc + shots[inv_shot[k]].img[(size_t)inv_col[k] * nz + iz];
    out[(size_t)ix * nz + iz] = acc;
}

}  // namespace sfd
