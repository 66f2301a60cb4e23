// seisfd3d.hip — 3D constant-density forward modeling on gfx950 (the
// reference's propagator_acoustic with genpar%twoD false: Acoustic_modeling_sep
// .f90:112-121, nbound 4).
//
// One shot at a time: the shot's window and its padded model live in HBM,
// five rotating wavefield levels.  Each time step runs, in the reference's
// order (Propagator.f90:168-243): receiver extraction on u2, the 25-point
// stencil u3 = v^2 ((lap_x + lap_z) + lap_y), the source's 7x7x7 sinc
// injection, the 5-level time update, and Boundary_3d's Higdon bands — z
// bands per (x, y) line, then x bands per (z, y) line, then y bands per (z, x)
// line, each line a sequential recursion outwards (Boundary.f90:89-184).  The
// stencil and update are one thread per cell with z the fastest (coalesced)
// index; the 3D path is a correctness path (seven launches a step), not tuned.
// Compiled with -ffp-contract=off: the reference's FP32 rounding bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/seisfd.h"
#include "common.h"

void sfd_host_mksinc(float *sinc, int lsinc, float shift);
void sfd_host_taper(float *t, int n);
void second_difference(std::vector<float> &w, float dt);

namespace sfd3 {

struct Geo {
    int NZ, NX, NY;   // padded dims (FD region + 4-cell halo per side)
    int LDZ;          // leading dimension (NZ rounded up to 32)
    size_t plane;     // LDZ * NX
    __host__ __device__ size_t at(int iz, int ix, int iy) const {
        return (size_t)iy * plane + (size_t)ix * LDZ + (size_t)iz;
    }
};

struct Coef3 { float c[15]; };   // c0..c4 per axis: z, x, y

// Boundary.f90:66-87 Higdon_coeff
__device__ __forceinline__ void higdon_coeff3(float nu, float *c) {
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

// mod_window_pad + model_pad (3D): the window's value at the clamped cell;
// v2 = v * v.  colsrc / rowsrc: model column / line (0-based) of each window
// column / line.
__global__ void pad3_kernel(const float *__restrict__ big, int nz, int nx, const int *__restrict__ colsrc, int nxw,
                            const int *__restrict__ linesrc, int nyw, int ntaper, Geo G, float *__restrict__ vel,
                            float *__restrict__ v2) {
    const int iz = blockIdx.x * blockDim.x + threadIdx.x, ix = blockIdx.y, iy = blockIdx.z;
    if (iz >= G.NZ) return;
    int wz = iz - 4 - ntaper, wx = ix - 4 - ntaper, wy = iy - 4 - ntaper;
    wz = wz < 0 ? 0 : (wz >= nz ? nz - 1 : wz);
    wx = wx < 0 ? 0 : (wx >= nxw ? nxw - 1 : wx);
    wy = wy < 0 ? 0 : (wy >= nyw ? nyw - 1 : wy);
    const float v = (float)0 + big[(size_t)wz + (size_t)colsrc[wx] * nz + (size_t)linesrc[wy] * nz * nx];
    const size_t q = G.at(iz, ix, iy);
    vel[q] = v;
    if (v2) v2[q] = v * v;
}

// Interpolate.f90:44-130 in 3D, one pass: out = the 9-point sinc (shift 0.5)
// along one axis of the padded box, each line extended by its end values
// (clamped indices), summed l = -4..4; the last pass stores 1 / rho2
// (allocate_sxx3d_syy3d_szz3d_delp3d, FDderivatives.f90:393).
struct Sinc9 { float w[9]; };
__global__ void rho_half3_kernel(const float *__restrict__ in, float *__restrict__ out, Geo G, Sinc9 S, int axis,
                                 int invert) {
    const int iz = blockIdx.x * blockDim.x + threadIdx.x, ix = blockIdx.y, iy = blockIdx.z;
    if (iz >= G.NZ) return;
    const int n = axis == 0 ? G.NZ : axis == 1 ? G.NX : G.NY;
    const int c = axis == 0 ? iz : axis == 1 ? ix : iy;
    float f = 0.f;
    for (int l = -4; l <= 4; l++) {
        int m = c + l;
        m = m < 0 ? 0 : (m >= n ? n - 1 : m);
        const float t = axis == 0 ? in[G.at(m, ix, iy)] : axis == 1 ? in[G.at(iz, m, iy)] : in[G.at(iz, ix, m)];
        f = f + S.w[4 + l] * t;
    }
    out[G.at(iz, ix, iy)] = invert ? 1.f / f : f;
}

// FDderivatives.f90:263-382, first half: the staggered gradients s3dzz,
// s3dxx, s3dyy times 1/rho2 on the FD region, and first-order gradients
// times the face's 1/rho2 in the three operator-padding layers of each face
// (top/bottom over FD (x, y), left/right over FD (z, y), front/back over FD
// (z, x)).  The outermost layer, edges and corners are never written (they
// keep the zero of the shot's start).  One thread per cell.
__global__ void __launch_bounds__(256) rho_grad3_kernel(const float *__restrict__ u2, const float *__restrict__ delp,
                                                        float *__restrict__ sz, float *__restrict__ sx,
                                                        float *__restrict__ sy, Geo G, Coef3 C, float dzi, float dxi,
                                                        float dyi) {
    const int iz = blockIdx.x * 64 + (threadIdx.x & 63), ix = blockIdx.y * 4 + (threadIdx.x >> 6),
              iy = blockIdx.z;
    if (iz >= G.NZ || ix >= G.NX) return;
    const bool fz = iz >= 4 && iz <= G.NZ - 5, fx = ix >= 4 && ix <= G.NX - 5, fy = iy >= 4 && iy <= G.NY - 5;
    const bool pz = (iz >= 1 && iz <= 3) || (iz >= G.NZ - 4 && iz <= G.NZ - 2);
    const bool px = (ix >= 1 && ix <= 3) || (ix >= G.NX - 4 && ix <= G.NX - 2);
    const bool py = (iy >= 1 && iy <= 3) || (iy >= G.NY - 4 && iy <= G.NY - 2);
    const size_t q = G.at(iz, ix, iy), X = G.LDZ, Y = G.plane;
    const float *c = C.c;
#define U(o) u2[q + (o)]
    if (fz && fx && fy) {
        const float d = delp[q];
        sz[q] = (c[1] * (U(1) - U(0)) + c[2] * (U(2) - U(-1)) + c[3] * (U(3) - U(-2)) + c[4] * (U(4) - U(-3))) * d;
        sx[q] = (c[6] * (U(X) - U(0)) + c[7] * (U(2 * X) - U(-(long)X)) + c[8] * (U(3 * X) - U(-2 * (long)X)) +
                 c[9] * (U(4 * X) - U(-3 * (long)X))) * d;
        sy[q] = (c[11] * (U(Y) - U(0)) + c[12] * (U(2 * Y) - U(-(long)Y)) + c[13] * (U(3 * Y) - U(-2 * (long)Y)) +
                 c[14] * (U(4 * Y) - U(-3 * (long)Y))) * d;
        return;
    }
    size_t e;   // the face cell whose 1/rho2 the padding takes
    if (pz && fx && fy) e = G.at(iz < 4 ? 4 : G.NZ - 5, ix, iy);
    else if (fz && px && fy) e = G.at(iz, ix < 4 ? 4 : G.NX - 5, iy);
    else if (fz && fx && py) e = G.at(iz, ix, iy < 4 ? 4 : G.NY - 5);
    else return;
    const float d = delp[e];
    sz[q] = dzi * (U(1) - U(0)) * d;
    sx[q] = dxi * (U(X) - U(0)) * d;
    sy[q] = dyi * (U(Y) - U(0)) * d;
#undef U
}

// Boundary.f90:36-60 Higdon tables: gz per (x, y) of the FD region from the
// top / bottom FD rows, gx per (z, y) from the left / right FD columns, gy
// per (z, x) from the front / back FD lines.
__global__ void higdon3_kernel(const float *__restrict__ vel, Geo G, float dt, float dz, float dx, float dy,
                               float *__restrict__ gz, float *__restrict__ gx, float *__restrict__ gy) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y, which = blockIdx.z;
    const int fz = G.NZ - 8, fx = G.NX - 8, fy = G.NY - 8;
    const float dzi = 1.f / dz, dxi = 1.f / dx, dyi = 1.f / dy;
    if (which == 0) {   // a = x, b = y
        if (a >= fx || b >= fy) return;
        float *g = gz + 16 * ((size_t)a + (size_t)b * fx);
        higdon_coeff3(vel[G.at(4, a + 4, b + 4)] * dt * dzi, g);
        higdon_coeff3(vel[G.at(G.NZ - 5, a + 4, b + 4)] * dt * dzi, g + 8);
    } else if (which == 1) {   // a = z, b = y
        if (a >= fz || b >= fy) return;
        float *g = gx + 16 * ((size_t)a + (size_t)b * fz);
        higdon_coeff3(vel[G.at(a + 4, 4, b + 4)] * dt * dxi, g);
        higdon_coeff3(vel[G.at(a + 4, G.NX - 5, b + 4)] * dt * dxi, g + 8);
    } else {   // a = z, b = x
        if (a >= fz || b >= fx) return;
        float *g = gy + 16 * ((size_t)a + (size_t)b * fz);
        higdon_coeff3(vel[G.at(a + 4, b + 4, 4)] * dt * dyi, g);
        higdon_coeff3(vel[G.at(a + 4, b + 4, G.NY - 5)] * dt * dyi, g + 8);
    }
}

// Injection_sinc -> Injection_source_sinc_xyz (Injection.f90:151-194,
// :274-346, shot_type 0) of one or more traces, as the fused step kernels
// apply it: the plan lists each footprint cell's contributions in trace
// order (the reference adds trace after trace); each is u3 += (((v2r s) wz)
// wx) wy, v2r the trace's v^2 (times rho) and the y weight the x sinc, as the
// reference computes them.
struct InjC { int tr; float wz, wx, wy; };

// One step of the constant-density scheme fused into one pass over the FD
// region: the 25-point stencil u3 = v^2 (tx + tz + ty) (FDderivatives.f90:
// 588-657 with the 3D forward grid) or, with F3_ADJ, the adjoint-style
// stencil of v^2 u2 (FD_2nd_3D_derivatives_scalar_adjoint_grid,
// FDderivatives.f90:780-819), the injection of the traces whose footprints
// hold the cell (Injection_sinc or Injection_LSRTM_sinc, in trace order),
// with F3_BORN Injection_Born's 2 t m S / v on the window
// (Imaging.f90:747-788), and the time update (FDderivatives.f90:960-982),
// in the reference's order of operations.  A workgroup owns a 64 (z) x 4 CPT (x) tile
// and walks yb planes along y: each lane keeps the stencil's operand (u2, or
// v^2 u2) of its row for its CPT columns and the 4 planes either side in
// registers (the y-derivative) and stages the plane in LDS with a 4-cell
// ring (the z- and x-derivatives); the next plane's ring, centre and update
// operands are loaded while the current plane is computed.  Cells inside the
// injection box look up their line in colptr (the injection cells per (x, y)
// line of the padded grid).
constexpr int F3_TZ = 64, F3_LZ = F3_TZ + 8;
struct Box3 { int z0, z1, x0, x1, y0, y1; };   // padded cells holding injections (empty: z0 > z1)
struct Born3 {                                  // Injection_Born's operands (F3_BORN)
    const float *m, *vel, *S, *tx, *ty;         // padded reflectivity and velocity, window snapshot, tapers
    int nz, nxw, nyw, i0;
};
template <int F3_CPT, bool F3_ADJ, bool F3_BORN>
__global__ void __launch_bounds__(256) step3_fused_kernel(
    const float *__restrict__ um, const float *__restrict__ u0, const float *__restrict__ u1,
    const float *__restrict__ u2, float *__restrict__ u3, const float *__restrict__ v2, Geo G, Coef3 C,
    float c12dt2, float div11, int yb, Box3 box, const int *__restrict__ colptr, const size_t *__restrict__ cell,
    const int *__restrict__ cptr, const InjC *__restrict__ con, const float *__restrict__ v2r,
    const float *__restrict__ traces, int nt, int it, Born3 born) {
    constexpr int TX = 4 * F3_CPT, LX = TX + 8, NRING = 8 * TX + 8 * F3_LZ, RPT = (NRING + 255) / 256;
    __shared__ float sh[LX * F3_LZ];   // [column][row], rows iz0-4 .. iz0+67, columns ix0-4 .. ix0+TX+3
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // Workgroups are dealt round-robin to the 8 XCDs, each with its own L2:
    // renumber them so every XCD takes a contiguous run of tiles (z fastest,
    // then x, then y), whose rings and planes its L2 then shares.
    const int nbx = (int)gridDim.x, nby = (int)gridDim.y;
    const int nb = nbx * nby * (int)gridDim.z;
    const int hw = (int)blockIdx.x + nbx * ((int)blockIdx.y + nby * (int)blockIdx.z);
    const int xq = nb / 8, xr = nb % 8, xcd = hw % 8;
    const int lb = xcd * xq + min(xcd, xr) + hw / 8;
    const int bx = lb % nbx, by = (lb / nbx) % nby, bz = lb / (nbx * nby);
    const int iz0 = 4 + bx * F3_TZ, ix0 = 4 + by * TX;
    const int iy0 = 4 + bz * yb, iy1 = min(iy0 + yb, G.NY - 4);
    const int iz = iz0 + lane, izl = min(iz, G.NZ - 1);
    const bool zok = iz <= G.NZ - 5;
    int ixc[F3_CPT];
    bool ok[F3_CPT];
#pragma unroll
    for (int c = 0; c < F3_CPT; c++) {
        const int ix = ix0 + F3_CPT * w + c;
        ok[c] = zok && ix <= G.NX - 5;
        ixc[c] = min(ix, G.NX - 1);
    }
    // the stencil's operand at a cell: u2, or v^2 u2 for the adjoint-style stencil
    auto opnd = [&](size_t a) -> float { return F3_ADJ ? v2[a] * u2[a] : u2[a]; };
    // this thread's ring cells: 8 rows x TX columns above and below, 8 columns x 72 rows beside
    int rl[RPT];
    size_t rg[RPT];
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        const int t = threadIdx.x + 256 * r;
        int col = 0, row = 0;
        if (t < 8 * TX) {
            col = 4 + t / 8;
            const int m = t % 8;
            row = m < 4 ? m : F3_TZ + m;
        } else if (t < NRING) {
            const int u = t - 8 * TX, cc = u / F3_LZ;
            col = cc < 4 ? cc : TX + cc;
            row = u % F3_LZ;
        }
        rl[r] = t < NRING ? col * F3_LZ + row : -1;
        rg[r] = G.at(min(iz0 - 4 + row, G.NZ - 1), min(ix0 - 4 + col, G.NX - 1), 0);
    }
    // prologue: the queue, then plane iy0's centre, ring and update operands
    // (vv: v^2 for the forward stencil's product, u2 itself for the adjoint's update)
    float q[F3_CPT][9], nq[F3_CPT], vv[F3_CPT], mm[F3_CPT], z0[F3_CPT], z1[F3_CPT], rv[RPT];
#pragma unroll
    for (int c = 0; c < F3_CPT; c++) {
#pragma unroll
        for (int m = 0; m < 8; m++) q[c][m] = opnd(G.at(izl, ixc[c], iy0 - 4 + m));
        nq[c] = opnd(G.at(izl, ixc[c], iy0 + 4));
        const size_t a = G.at(izl, ixc[c], iy0);
        vv[c] = F3_ADJ ? u2[a] : v2[a];
        mm[c] = um[a];
        z0[c] = u0[a];
        z1[c] = u1[a];
    }
#pragma unroll
    for (int r = 0; r < RPT; r++) rv[r] = rl[r] >= 0 ? opnd(rg[r] + (size_t)iy0 * G.plane) : 0.f;
    const float *c = C.c;
    const bool box_z = iz >= box.z0 && iz <= box.z1;
    for (int iy = iy0; iy < iy1; iy++) {
        // stage plane iy (loaded one iteration ahead)
#pragma unroll
        for (int k = 0; k < F3_CPT; k++) {
            q[k][8] = nq[k];
            sh[(4 + F3_CPT * w + k) * F3_LZ + 4 + lane] = q[k][4];
        }
#pragma unroll
        for (int r = 0; r < RPT; r++)
            if (rl[r] >= 0) sh[rl[r]] = rv[r];
        __syncthreads();
        // issue plane iy+1's loads; they are used in the next iteration
        float nv[F3_CPT], nm[F3_CPT], n0[F3_CPT], n1[F3_CPT];
        const bool more = iy + 1 < iy1;
        if (more) {
#pragma unroll
            for (int k = 0; k < F3_CPT; k++) {
                nq[k] = opnd(G.at(izl, ixc[k], iy + 5));
                const size_t a = G.at(izl, ixc[k], iy + 1);
                nv[k] = F3_ADJ ? u2[a] : v2[a];
                nm[k] = um[a];
                n0[k] = u0[a];
                n1[k] = u1[a];
            }
#pragma unroll
            for (int r = 0; r < RPT; r++)
                if (rl[r] >= 0) rv[r] = opnd(rg[r] + (size_t)(iy + 1) * G.plane);
        }
#pragma unroll
        for (int k = 0; k < F3_CPT; k++) {
            if (ok[k]) {
                const float *S = sh + (4 + F3_CPT * w + k) * F3_LZ + 4 + lane;
#define UZ(o) S[(o)]
#define UX(o) S[(o) * F3_LZ]
                const float tz = c[0] * (UZ(0)) + c[1] * (UZ(1) + UZ(-1)) + c[2] * (UZ(2) + UZ(-2)) +
                                 c[3] * (UZ(3) + UZ(-3)) + c[4] * (UZ(4) + UZ(-4));
                const float tx = c[5] * (UX(0)) + c[6] * (UX(1) + UX(-1)) + c[7] * (UX(2) + UX(-2)) +
                                 c[8] * (UX(3) + UX(-3)) + c[9] * (UX(4) + UX(-4));
#undef UZ
#undef UX
                const float ty = c[10] * (q[k][4]) + c[11] * (q[k][5] + q[k][3]) + c[12] * (q[k][6] + q[k][2]) +
                                 c[13] * (q[k][7] + q[k][1]) + c[14] * (q[k][8] + q[k][0]);
                const size_t a = G.at(iz, ixc[k], iy);
                float u = F3_ADJ ? tx + tz + ty : vv[k] * (tx + tz + ty);
                if (box_z && ixc[k] >= box.x0 && ixc[k] <= box.x1 && iy >= box.y0 && iy <= box.y1) {
                    const int l = iy * G.NX + ixc[k];
                    for (int e = colptr[l]; e < colptr[l + 1]; e++) {
                        if (cell[e] != a) continue;
                        for (int m = cptr[e]; m < cptr[e + 1]; m++) {
                            const InjC f = con[m];
                            u = u + v2r[f.tr] * traces[(size_t)f.tr * nt + (it - 1)] * f.wz * f.wx * f.wy;
                        }
                    }
                }
                if constexpr (F3_BORN) {
                    const int i = iz - born.i0, j = ixc[k] - born.i0, l = iy - born.i0;
                    if (i >= 0 && i < born.nz && j >= 0 && j < born.nxw && l >= 0 && l < born.nyw) {
                        const size_t wq = (size_t)i + (size_t)j * born.nz + (size_t)l * born.nz * born.nxw;
                        const float t = born.tx[j] * born.ty[l];
                        u = u + 2.f * t * born.m[a] * born.S[wq] / born.vel[a];
                    }
                }
                const float u2c = F3_ADJ ? vv[k] : q[k][4];
                u3[a] = (20.f * u2c - 6.f * z1[k] - 4.f * z0[k] + mm[k] + c12dt2 * u) * div11;
            }
#pragma unroll
            for (int m = 0; m < 8; m++) q[k][m] = q[k][m + 1];
            if (more) {
                vv[k] = nv[k];
                mm[k] = nm[k];
                z0[k] = n0[k];
                z1[k] = n1[k];
            }
        }
        __syncthreads();
    }
}

// The variable-density step after rho_grad3_kernel, fused into one pass over
// the FD region: the staggered divergence u3 = v^2 rho ((tx + ty) + tz) of
// the gradients sz, sx, sy (FDderivatives.f90:263-382), the injection
// (Injection.f90:151-194, in trace order) and the time update
// (FDderivatives.f90:960-982), in the reference's order of operations.  Same walk as step3_fused_kernel: a 64 (z) x
// 4 CPT (x) tile walks yb planes along y, sy of the planes iy-4 .. iy+3 in
// registers, sz of the plane in LDS with 4 rows above and below, sx with 4
// columns each side; the next plane's loads are issued before the current
// plane is computed.
template <int F3_CPT>
__global__ void __launch_bounds__(256) step3_rho_fused_kernel(
    const float *__restrict__ um, const float *__restrict__ u0, const float *__restrict__ u1,
    const float *__restrict__ u2, float *__restrict__ u3, const float *__restrict__ sz, const float *__restrict__ sx,
    const float *__restrict__ sy, const float *__restrict__ v2, const float *__restrict__ rho, Geo G, Coef3 C,
    float c12dt2, float div11, int yb, Box3 box, const int *__restrict__ colptr, const size_t *__restrict__ cell,
    const int *__restrict__ cptr, const InjC *__restrict__ con, const float *__restrict__ v2r,
    const float *__restrict__ traces, int nt, int it) {
    constexpr int TX = 4 * F3_CPT, ZL = F3_LZ, XL = TX + 8;
    constexpr int NZR = 8 * TX, NXR = 8 * F3_TZ, NRING = NZR + NXR, RPT = (NRING + 255) / 256;
    __shared__ float shz[TX * ZL];        // sz: [column][row], rows iz0-4 .. iz0+67
    __shared__ float shx[XL * F3_TZ];     // sx: [column][row], columns ix0-4 .. ix0+TX+3
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nbx = (int)gridDim.x, nby = (int)gridDim.y;
    const int nb = nbx * nby * (int)gridDim.z;
    const int hw = (int)blockIdx.x + nbx * ((int)blockIdx.y + nby * (int)blockIdx.z);
    const int xq = nb / 8, xr = nb % 8, xcd = hw % 8;
    const int lb = xcd * xq + min(xcd, xr) + hw / 8;
    const int bx = lb % nbx, by = (lb / nbx) % nby, bz = lb / (nbx * nby);
    const int iz0 = 4 + bx * F3_TZ, ix0 = 4 + by * TX;
    const int iy0 = 4 + bz * yb, iy1 = min(iy0 + yb, G.NY - 4);
    const int iz = iz0 + lane, izl = min(iz, G.NZ - 1);
    const bool zok = iz <= G.NZ - 5;
    int ixc[F3_CPT];
    bool ok[F3_CPT];
#pragma unroll
    for (int c = 0; c < F3_CPT; c++) {
        const int ix = ix0 + F3_CPT * w + c;
        ok[c] = zok && ix <= G.NX - 5;
        ixc[c] = min(ix, G.NX - 1);
    }
    // ring cells: sz 8 rows x TX columns above and below, sx 8 columns x 64 rows beside
    int rl[RPT];
    bool rz[RPT];
    size_t rg[RPT];
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        const int t = threadIdx.x + 256 * r;
        int col = 0, row = 0;
        rz[r] = t < NZR;
        if (t < NZR) {
            col = t / 8;
            const int m = t % 8;
            row = m < 4 ? m : F3_TZ + m;   // LDS rows 0..3 and 68..71
            rl[r] = col * ZL + row;
            rg[r] = G.at(min(iz0 - 4 + row, G.NZ - 1), min(ix0 + col, G.NX - 1), 0);
        } else if (t < NRING) {
            const int u = t - NZR, cc = u / F3_TZ;
            col = cc < 4 ? cc : TX + cc;    // LDS columns 0..3 and TX+4..TX+7
            row = u % F3_TZ;
            rl[r] = col * F3_TZ + row;
            rg[r] = G.at(min(iz0 + row, G.NZ - 1), min(ix0 - 4 + col, G.NX - 1), 0);
        } else {
            rl[r] = -1;
            rg[r] = 0;
        }
    }
    float q[F3_CPT][8], nq[F3_CPT], cz[F3_CPT], cx[F3_CPT], ncz[F3_CPT], ncx[F3_CPT];
    float vr[F3_CPT], uu[F3_CPT], mm[F3_CPT], z0[F3_CPT], z1[F3_CPT], rv[RPT];
#pragma unroll
    for (int c = 0; c < F3_CPT; c++) {
#pragma unroll
        for (int m = 0; m < 7; m++) q[c][m] = sy[G.at(izl, ixc[c], iy0 - 4 + m)];
        nq[c] = sy[G.at(izl, ixc[c], iy0 + 3)];
        const size_t a = G.at(izl, ixc[c], iy0);
        cz[c] = sz[a];
        cx[c] = sx[a];
        vr[c] = v2[a] * rho[a];
        uu[c] = u2[a];
        mm[c] = um[a];
        z0[c] = u0[a];
        z1[c] = u1[a];
    }
#pragma unroll
    for (int r = 0; r < RPT; r++)
        rv[r] = rl[r] >= 0 ? (rz[r] ? sz : sx)[rg[r] + (size_t)iy0 * G.plane] : 0.f;
    const float *c = C.c;
    const bool box_z = iz >= box.z0 && iz <= box.z1;
    for (int iy = iy0; iy < iy1; iy++) {
#pragma unroll
        for (int k = 0; k < F3_CPT; k++) {
            q[k][7] = nq[k];
            shz[(F3_CPT * w + k) * ZL + 4 + lane] = cz[k];
            shx[(4 + F3_CPT * w + k) * F3_TZ + lane] = cx[k];
        }
#pragma unroll
        for (int r = 0; r < RPT; r++)
            if (rl[r] >= 0) (rz[r] ? shz : shx)[rl[r]] = rv[r];
        __syncthreads();
        float nvr[F3_CPT], nuu[F3_CPT], nm[F3_CPT], n0[F3_CPT], n1[F3_CPT];
        const bool more = iy + 1 < iy1;
        if (more) {
#pragma unroll
            for (int k = 0; k < F3_CPT; k++) {
                nq[k] = sy[G.at(izl, ixc[k], iy + 4)];
                const size_t a = G.at(izl, ixc[k], iy + 1);
                ncz[k] = sz[a];
                ncx[k] = sx[a];
                nvr[k] = v2[a] * rho[a];
                nuu[k] = u2[a];
                nm[k] = um[a];
                n0[k] = u0[a];
                n1[k] = u1[a];
            }
#pragma unroll
            for (int r = 0; r < RPT; r++)
                if (rl[r] >= 0) rv[r] = (rz[r] ? sz : sx)[rg[r] + (size_t)(iy + 1) * G.plane];
        }
#pragma unroll
        for (int k = 0; k < F3_CPT; k++) {
            if (ok[k]) {
                const float *Z = shz + (F3_CPT * w + k) * ZL + 4 + lane;       // sz at row iz
                const float *X = shx + (4 + F3_CPT * w + k) * F3_TZ + lane;    // sx at column ix
                const float tz = c[1] * (Z[0] - Z[-1]) + c[2] * (Z[1] - Z[-2]) + c[3] * (Z[2] - Z[-3]) +
                                 c[4] * (Z[3] - Z[-4]);
                const float tx = c[6] * (X[0] - X[-F3_TZ]) + c[7] * (X[F3_TZ] - X[-2 * F3_TZ]) +
                                 c[8] * (X[2 * F3_TZ] - X[-3 * F3_TZ]) + c[9] * (X[3 * F3_TZ] - X[-4 * F3_TZ]);
                const float ty = c[11] * (q[k][4] - q[k][3]) + c[12] * (q[k][5] - q[k][2]) +
                                 c[13] * (q[k][6] - q[k][1]) + c[14] * (q[k][7] - q[k][0]);
                const size_t a = G.at(iz, ixc[k], iy);
                float u = vr[k] * (tx + ty + tz);
                if (box_z && ixc[k] >= box.x0 && ixc[k] <= box.x1 && iy >= box.y0 && iy <= box.y1) {
                    const int l = iy * G.NX + ixc[k];
                    for (int e = colptr[l]; e < colptr[l + 1]; e++) {
                        if (cell[e] != a) continue;
                        for (int m = cptr[e]; m < cptr[e + 1]; m++) {
                            const InjC f = con[m];
                            u = u + v2r[f.tr] * traces[(size_t)f.tr * nt + (it - 1)] * f.wz * f.wx * f.wy;
                        }
                    }
                }
                u3[a] = (20.f * uu[k] - 6.f * z1[k] - 4.f * z0[k] + mm[k] + c12dt2 * u) * div11;
            }
#pragma unroll
            for (int m = 0; m < 7; m++) q[k][m] = q[k][m + 1];
            if (more) {
                cz[k] = ncz[k];
                cx[k] = ncx[k];
                vr[k] = nvr[k];
                uu[k] = nuu[k];
                mm[k] = nm[k];
                z0[k] = n0[k];
                z1[k] = n1[k];
            }
        }
        __syncthreads();
    }
}

// Extraction_wavefield (Extraction.f90:562-601): u2 on the window (rows 1..nz,
// window columns and lines) into snapshot slot s.
__global__ void snap3_kernel(const float *__restrict__ u2, Geo G, int nz, int nxw, int i0,
                             float *__restrict__ snap) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= nz) return;
    snap[(size_t)i + (size_t)j * nz + (size_t)k * nz * nxw] = u2[G.at(i + i0, j + i0, k + i0)];
}

// Imaging_condition_sourceonly_from_memory (Imaging.f90:228-275): image +=
// u S t, illum += S S t with t = taperx(j) tapery(k), per window cell.
__global__ void image3_kernel(const float *__restrict__ u2, Geo G, int nz, int nxw, int i0,
                              const float *__restrict__ snap, const float *__restrict__ tx,
                              const float *__restrict__ ty, float *__restrict__ img, float *__restrict__ ill) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= nz) return;
    const size_t q = (size_t)i + (size_t)j * nz + (size_t)k * nz * nxw;
    const float t = tx[j] * ty[k], S = snap[q];
    img[q] = img[q] + u2[G.at(i + i0, j + i0, k + i0)] * S * t;
    ill[q] = ill[q] + S * S * t;
}

// Imaging_condition_AFWI_from_memory (Imaging.f90:371-422) in 3D: image +=
// u ((S- - 2 S) + S+) / (dt2 v^3) t, illum += S S t, per window cell.
__global__ void afwi3_image_kernel(const float *__restrict__ u2, const float *__restrict__ vel, Geo G, int nz, int nxw,
                                   int i0, const float *__restrict__ S, const float *__restrict__ Sm,
                                   const float *__restrict__ Sp, const float *__restrict__ tx,
                                   const float *__restrict__ ty, float dt2, float *__restrict__ img,
                                   float *__restrict__ ill) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= nz) return;
    const size_t q = (size_t)i + (size_t)j * nz + (size_t)k * nz * nxw, a = G.at(i + i0, j + i0, k + i0);
    const float t = tx[j] * ty[k], v = vel[a];
    const float tmp = (Sm[q] - 2.f * S[q] + Sp[q]) / (dt2 * (v * v * v));
    img[q] = img[q] + u2[a] * tmp * t;
    ill[q] = ill[q] + S[q] * S[q] * t;
}

// Imaging_condition_AFWI_RHOVP_from_memory (Imaging.f90:424-512) in 3D: with
// ip = min(nz, i+1), jp = min(j+1, nxw), kp = min(k+dk, nyw) (dk = 0 when
// nyw = 1), im = ip-1, jm = jp-1, km = kp-dk:
//   tmp  = dzi du_z dzi dS_z + dxi du_x dxi dS_x + dyi du_y dyi dS_y (one-sided)
//   tmp1 = u / dt2 (S- - 2 S + S+)
//   image += 2 tmp1 t / (v^3 rho), image2 += tmp t / rho^2 + tmp1 t / (rho^2 v^2)
__global__ void rhovp3_image_kernel(const float *__restrict__ u2, const float *__restrict__ vel,
                                    const float *__restrict__ rho, Geo G, int nz, int nxw, int nyw, int i0,
                                    const float *__restrict__ S, const float *__restrict__ Sm,
                                    const float *__restrict__ Sp, const float *__restrict__ tx,
                                    const float *__restrict__ ty, float dt2, float dzi, float dxi, float dyi,
                                    float *__restrict__ img, float *__restrict__ img2, float *__restrict__ ill) {
    const int i0w = blockIdx.x * blockDim.x + threadIdx.x, j0w = blockIdx.y, k0w = blockIdx.z;
    if (i0w >= nz) return;
    const int i = i0w + 1, j = j0w + 1, k = k0w + 1, dk = nyw == 1 ? 0 : 1;
    const int ip = i + 1 < nz ? i + 1 : nz, jp = j + 1 < nxw ? j + 1 : nxw, kp = k + dk < nyw ? k + dk : nyw;
    const int im = ip - 1, jm = jp - 1, km = kp - dk;
    auto U = [&](int a, int b, int c) { return u2[G.at(a - 1 + i0, b - 1 + i0, c - 1 + i0)]; };
    auto W = [&](int a, int b, int c) { return S[(size_t)(a - 1) + (size_t)(b - 1) * nz + (size_t)(c - 1) * nz * nxw]; };
    float tmp = dzi * (U(ip, jm, km) - U(im, jm, km)) * dzi * (W(ip, jm, km) - W(im, jm, km));
    tmp = tmp + dxi * (U(im, jp, km) - U(im, jm, km)) * dxi * (W(im, jp, km) - W(im, jm, km));
    tmp = tmp + dyi * (U(im, jm, kp) - U(im, jm, km)) * dyi * (W(im, jm, kp) - W(im, jm, km));
    const size_t q = (size_t)i0w + (size_t)j0w * nz + (size_t)k0w * nz * nxw, a = G.at(i0w + i0, j0w + i0, k0w + i0);
    const float tmp1 = u2[a] * (1.f / dt2) * (Sm[q] - 2.f * S[q] + Sp[q]);
    const float t = tx[j0w] * ty[k0w], v = vel[a], r = rho[a];
    img[q] = img[q] + 2.f * tmp1 * t / (v * v * v * r);
    img2[q] = img2[q] + tmp * t / (r * r) + tmp1 * t / (r * r * (v * v));
    ill[q] = ill[q] + S[q] * S[q] * t;
}

// Imaging_condition_LSRTM_sourceonly_from_memory (Imaging.f90:698-745) in 3D:
// image += 2 t u S / v, illum += S S t.
__global__ void lsrtm3_image_kernel(const float *__restrict__ u2, const float *__restrict__ vel, Geo G, int nz,
                                    int nxw, int i0, const float *__restrict__ S, const float *__restrict__ tx,
                                    const float *__restrict__ ty, float *__restrict__ img, float *__restrict__ ill) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= nz) return;
    const size_t q = (size_t)i + (size_t)j * nz + (size_t)k * nz * nxw, a = G.at(i + i0, j + i0, k + i0);
    const float t = tx[j] * ty[k];
    img[q] = img[q] + 2.f * t * u2[a] * S[q] / vel[a];
    ill[q] = ill[q] + S[q] * S[q] * t;
}

// Higdon_factor (Boundary.f90:366-447) at cell q along a line of stride s
// (s < 0: the band lies at the low end and the recursion reads q - s, q - 2s)
__device__ __forceinline__ void hfac(const float *g, float *U3, const float *U2, const float *U1, size_t q, long s) {
    const float f = g[0] * U3[q + s] + g[1] * U3[q + 2 * s] + g[2] * U2[q] + g[3] * U2[q + s] +
                    g[4] * U2[q + 2 * s] + g[5] * U1[q] + g[6] * U1[q + s] + g[7] * U1[q + 2 * s];
    U3[q] = -f;
}

// z bands of every (x, y) line of the FD region: rows nmin1+3 .. nmin1-4
// (padded 7 .. 0), then nmax1-3 .. nmax1+4, with the ten rows of each end of
// a line staged in LDS (a line's rows are contiguous): 8 lines per 256-thread
// block, 32 lanes per line, one lane per end for the recursion, in the order
// and arithmetic of hfac.
__global__ void __launch_bounds__(256) zband3s_kernel(float *__restrict__ U3, const float *__restrict__ U2,
                                                      const float *__restrict__ U1, const float *__restrict__ gz,
                                                      Geo G) {
    __shared__ float s3[8][20], s2[8][20], s1[8][20];
    const int col = threadIdx.x >> 5, l = threadIdx.x & 31;
    const int a = (int)blockIdx.x * 8 + col, b = blockIdx.y;   // FD x, FD y
    const bool on = a < G.NX - 8;
    const int slot = l < 10 ? l : (l >= 16 && l < 26 ? l - 6 : -1);
    const int row = l < 10 ? l : G.NZ - 26 + l;                 // rows 0..9, NZ-10..NZ-1
    size_t q = 0;
    if (on && slot >= 0) {
        q = G.at(row, a + 4, b + 4);
        s3[col][slot] = U3[q];
        s2[col][slot] = U2[q];
        s1[col][slot] = U1[q];
    }
    __syncthreads();
    if (on && (l == 0 || l == 16)) {
        const float *g = gz + 16 * ((size_t)a + (size_t)b * (G.NX - 8)) + (l == 0 ? 0 : 8);
        float *S3 = s3[col];
        const float *S2 = s2[col], *S1 = s1[col];
        if (l == 0) {
            for (int i = 7; i >= 0; i--)
                S3[i] = -(g[0] * S3[i + 1] + g[1] * S3[i + 2] + g[2] * S2[i] + g[3] * S2[i + 1] + g[4] * S2[i + 2] +
                          g[5] * S1[i] + g[6] * S1[i + 1] + g[7] * S1[i + 2]);
        } else {
            for (int i = 12; i <= 19; i++)
                S3[i] = -(g[0] * S3[i - 1] + g[1] * S3[i - 2] + g[2] * S2[i] + g[3] * S2[i - 1] + g[4] * S2[i - 2] +
                          g[5] * S1[i] + g[6] * S1[i - 1] + g[7] * S1[i - 2]);
        }
    }
    __syncthreads();
    if (on && ((slot >= 0 && slot < 8) || slot >= 12)) U3[q] = s3[col][slot];
}
// x bands of every (z, y) line with z in the FD rows and y in the FD lines
__global__ void xband3_kernel(float *__restrict__ U3, const float *__restrict__ U2, const float *__restrict__ U1,
                              const float *__restrict__ gx, Geo G) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;   // FD z, FD y
    if (a >= G.NZ - 8) return;
    const float *g = gx + 16 * ((size_t)a + (size_t)b * (G.NZ - 8));
    const long s = (long)G.LDZ;
    for (int j = 7; j >= 0; j--) hfac(g, U3, U2, U1, G.at(a + 4, j, b + 4), s);
    for (int j = G.NX - 8; j <= G.NX - 1; j++) hfac(g + 8, U3, U2, U1, G.at(a + 4, j, b + 4), -s);
}
// y bands of every (z, x) line with z in the FD rows and x in the FD columns
__global__ void yband3_kernel(float *__restrict__ U3, const float *__restrict__ U2, const float *__restrict__ U1,
                              const float *__restrict__ gy, Geo G) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;   // FD z, FD x
    if (a >= G.NZ - 8) return;
    const float *g = gy + 16 * ((size_t)a + (size_t)b * (G.NZ - 8));
    const long s = (long)G.plane;
    for (int k = 7; k >= 0; k--) hfac(g, U3, U2, U1, G.at(a + 4, b + 4, k), s);
    for (int k = G.NY - 8; k <= G.NY - 1; k++) hfac(g + 8, U3, U2, U1, G.at(a + 4, b + 4, k), -s);
}

// Receivers (Extraction.f90:71-162,217-246 -> :276-337), one thread each:
// nearest, or the 7x7x7 sinc summed k, j, i in order with the y weights from
// the x sinc; gain per receiver (v for SINC_V2 in 3D, 1/(dz dx dy), 1).
struct Rec3 { int iz, ix, iy; float gain; float wz[7], wx[7]; };
__global__ void extract3_kernel(const float *__restrict__ u2, Geo G, const Rec3 *__restrict__ R, int ntr,
                                int nearest, float *__restrict__ traces, int nt, int it) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= ntr) return;
    const Rec3 &e = R[r];
    float *tr = traces + (size_t)r * nt + (it - 1);
    if (nearest) {
        *tr = *tr + u2[G.at(e.iz, e.ix, e.iy)];
        return;
    }
    float acc = *tr;
    for (int k = -3; k <= 3; k++)
        for (int j = -3; j <= 3; j++)
            for (int i = -3; i <= 3; i++)
                acc = acc + e.gain * u2[G.at(e.iz + i, e.ix + j, e.iy + k)] * e.wz[3 + i] * e.wx[3 + j] * e.wx[3 + k];
    *tr = acc;
}

// The sinc receivers with one wavefront each: the lanes form the 343
// products of extract3_kernel (same factors, same order) into LDS, then lane
// 0 adds them to the trace sample k, j, i in order, as extract3_kernel does.
__global__ void __launch_bounds__(256) extract3w_kernel(const float *__restrict__ u2, Geo G,
                                                        const Rec3 *__restrict__ R, int ntr,
                                                        float *__restrict__ traces, int nt, int it) {
    __shared__ float p[4][343];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, r = (int)blockIdx.x * 4 + w;
    if (r < ntr) {
        const Rec3 &e = R[r];
        for (int t = lane; t < 343; t += 64) {
            const int i = t % 7 - 3, j = (t / 7) % 7 - 3, k = t / 49 - 3;
            p[w][t] = e.gain * u2[G.at(e.iz + i, e.ix + j, e.iy + k)] * e.wz[3 + i] * e.wx[3 + j] * e.wx[3 + k];
        }
    }
    __syncthreads();
    if (r < ntr && lane == 0) {
        float *tr = traces + (size_t)r * nt + (it - 1);
        float acc = *tr;
        for (int t = 0; t < 343; t++) acc = acc + p[w][t];
        *tr = acc;
    }
}

}  // namespace sfd3

using namespace sfd3;

static void extract3(hipStream_t st, const Geo &G, const float *u2, const Rec3 *rec, int ntr, int nearest,
                     float *trc, int nt, int it) {
    if (nearest)
        hipLaunchKernelGGL(extract3_kernel, dim3((ntr + 63) / 64), dim3(64), 0, st, u2, G, rec, ntr, 1, trc, nt, it);
    else
        hipLaunchKernelGGL(extract3w_kernel, dim3((ntr + 3) / 4), dim3(256), 0, st, u2, G, rec, ntr, trc, nt, it);
}

struct sfd_ctx3d {
    sfd_grid3d g;
    hipStream_t st = nullptr;
    std::vector<float> vel_h, velb_h, rho_h;   // velb_h: RTM's backward model; rho_h empty: constant density
    float *d_vel = nullptr, *d_velb = nullptr, *d_rho = nullptr;
    std::string err;
    void *mem = nullptr;
    size_t cap = 0;
    ~sfd_ctx3d() {
        if (mem) (void)hipFree(mem);
        if (d_vel) (void)hipFree(d_vel);
        if (d_velb) (void)hipFree(d_velb);
        if (d_rho) (void)hipFree(d_rho);
        if (st) (void)hipStreamDestroy(st);
    }
    char *reserve(size_t n) {
        if (n > cap) {
            if (mem) (void)hipFree(mem);
            mem = nullptr;
            cap = 0;
            if (hipMalloc(&mem, n) != hipSuccess) throw EngineError(SFD_ENOMEM, "device allocation failed (3D shot)");
            cap = n;
        }
        return (char *)mem;
    }
};

static int nint3(float x) { return (int)roundf(x); }
static void find_i_d3(int *ic, float *dc, float c, float o, float d) {
    *ic = nint3((c - o) / d) + 1;
    *dc = c - (o + (float)(*ic - 1) * d);
}

// extract_coord_source_receiver_patch (ExtractPadModel.f90:19-85) and
// nxw / nyw (:183-186)
static void window3(const sfd_grid3d &g, const sfd_shot3d &s, float *oxw, int *nxw, float *oyw, int *nyw) {
    float gx1 = s.rx[0], gxn = s.rx[0], gy1 = s.ry[0], gyn = s.ry[0];
    for (int i = 1; i < s.ntr; i++) {
        gx1 = std::min(gx1, s.rx[i]);
        gxn = std::max(gxn, s.rx[i]);
        gy1 = std::min(gy1, s.ry[i]);
        gyn = std::max(gyn, s.ry[i]);
    }
    const float minx = std::min(s.sx, gx1), miny = std::min(s.sy, gy1);
    const float maxx = std::max(s.sx, gxn), maxy = std::max(s.sy, gyn);
    const float mx1 = (gx1 + s.sx) / 2, mxn = (gxn + s.sx) / 2, my1 = (gy1 + s.sy) / 2, myn = (gyn + s.sy) / 2;
    const float cx = mx1 + fabsf(mxn - mx1) / 2, cy = my1 + fabsf(myn - my1) / 2;
    *oxw = std::min(minx, cx - g.aperture_x / 2);
    const float endx = std::max(maxx, cx + g.aperture_x / 2);
    *oyw = std::min(miny, cy - g.aperture_y / 2);
    const float endy = std::max(maxy, cy + g.aperture_y / 2);
    *nxw = nint3((endx - *oxw) / g.dx) + 1;
    *nyw = nint3((endy - *oyw) / g.dy) + 1;
}

static void check3(const sfd_grid3d &g) {
    if (g.nz < 2 || g.nx < 1 || g.ny < 1 || !(g.dz > 0) || !(g.dx > 0) || !(g.dy > 0) || !(g.dt > 0))
        throw EngineError(SFD_EINVAL, "bad 3D grid dimensions or spacing");
    if (g.ntaper < 10) throw EngineError(SFD_EINVAL, "ntaper must be >= 10");
    if (g.lsinc != 7) throw EngineError(SFD_EUNSUPPORTED, "only lsinc=7 is implemented");
    if (!(g.fmax >= 0.f) || !std::isfinite(g.fmax)) throw EngineError(SFD_EINVAL, "fmax must be >= 0 (0: 30 Hz)");
}

// ReadSouVelRho.f90:603-752 vel_check in 3D: Courant stop 0.35, sampling stop
static void vel_check3(const sfd_grid3d &g, const float *v, size_t n) {
    float vmax = v[0], vmin = v[0];
    for (size_t i = 1; i < n; i++) {
        vmax = std::max(vmax, v[i]);
        vmin = std::min(vmin, v[i]);
    }
    const float stab = vmax * g.dt / std::min(g.dz, std::min(g.dx, g.dy));
    char b[200];
    if (stab > 0.35f) {
        snprintf(b, sizeof b, "Courant number %.6g > 0.35: unstable in 3D", stab);
        throw EngineError(SFD_EUNSTABLE, b);
    }
    const float fmax = g.fmax > 0.f ? g.fmax : 30.f;
    const float samplerate = vmin / fmax / std::max(g.dz, std::max(g.dx, g.dy));
    if (samplerate < 2.0f) {
        snprintf(b, sizeof b, "wavefield not adequately sampled: %.6g cells per shortest wavelength", samplerate);
        throw EngineError(SFD_EALIASED, b);
    }
}

// One shot's run on the device: window, padded geometry, buffers, tables.
struct Run3 {
    const sfd_grid3d *g = nullptr;
    Geo G;
    float oxw = 0, oyw = 0;
    int nxw = 0, nyw = 0, i0 = 0, fz = 0, fx = 0, fy = 0;
    bool rho = false;
    std::vector<int> colsrc, linesrc;
    char *base = nullptr;
    float *u[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    float *vel = nullptr, *v2 = nullptr, *gz = nullptr, *gx = nullptr, *gy = nullptr;
    float *rhop = nullptr, *delp = nullptr, *tmp = nullptr, *sg[3] = {nullptr, nullptr, nullptr};
    const int *dcol = nullptr, *dlin = nullptr;
    Coef3 C;
    size_t nf = 0;
    float dzi = 0, dxi = 0, dyi = 0;
    // padded model value at a window cell (1-based, clamped as model_pad replicates edges)
    float at(const std::vector<float> &m, int icz, int icx, int icy) const {
        const int wz = std::min(std::max(icz, 1), g->nz) - 1, wx = std::min(std::max(icx, 1), nxw) - 1,
                  wy = std::min(std::max(icy, 1), nyw) - 1;
        return m[(size_t)wz + (size_t)colsrc[wx] * g->nz + (size_t)linesrc[wy] * g->nz * g->nx];
    }
};

// A host injection plan: footprint cells and their contributions in trace order.
struct InjPlan {
    std::vector<size_t> cell;   // footprint cells, ascending
    std::vector<int> cptr;      // contributions of cell k: con[cptr[k] .. cptr[k+1])
    std::vector<InjC> con;
    std::vector<float> v2r;
    std::vector<int> colptr;    // cells of (x, y) line l = iy NX + ix: cell[colptr[l] .. colptr[l+1])
};

// Injection_sinc's footprints (Injection.f90:181-187, :274-346): the 7x7x7
// cells of every trace, with v2r = vel^2 (times rho) at the trace's cell of
// the model `vm`; cells must lie in the FD region.
static void plan_inj3(const Run3 &R, const std::vector<float> &vm, const std::vector<float> &rh, int n,
                      const float *z, const float *x, const float *y, InjPlan &P) {
    const sfd_grid3d &g = *R.g;
    std::vector<std::pair<size_t, InjC>> all;
    all.reserve((size_t)n * 343);
    P.v2r.assign(std::max(n, 1), 0.f);
    for (int r = 0; r < n; r++) {
        int ic[3];
        float dc[3], wz[7], wx[7];
        find_i_d3(&ic[0], &dc[0], z[r], g.oz, g.dz);
        find_i_d3(&ic[1], &dc[1], x[r], R.oxw, g.dx);
        find_i_d3(&ic[2], &dc[2], y[r], R.oyw, g.dy);
        const int pz = ic[0] - 1 + R.i0, px = ic[1] - 1 + R.i0, py = ic[2] - 1 + R.i0;
        if (pz - 3 < 4 || pz + 3 > R.G.NZ - 5 || px - 3 < 4 || px + 3 > R.G.NX - 5 || py - 3 < 4 ||
            py + 3 > R.G.NY - 5)
            throw EngineError(SFD_EINVAL, "injection footprint leaves the FD region");
        const float v = R.at(vm, ic[0], ic[1], ic[2]);
        P.v2r[r] = v * v;
        if (R.rho) P.v2r[r] = P.v2r[r] * R.at(rh, ic[0], ic[1], ic[2]);
        sfd_host_mksinc(wz, 7, dc[0] / g.dz);
        sfd_host_mksinc(wx, 7, dc[1] / g.dx);
        for (int k = -3; k <= 3; k++)
            for (int j = -3; j <= 3; j++)
                for (int i = -3; i <= 3; i++)
                    all.push_back({R.G.at(pz + i, px + j, py + k), InjC{r, wz[3 + i], wx[3 + j], wx[3 + k]}});
    }
    std::stable_sort(all.begin(), all.end(),
                     [](const std::pair<size_t, InjC> &a, const std::pair<size_t, InjC> &b) { return a.first < b.first; });
    P.cell.clear();
    P.cptr.assign(1, 0);
    P.con.clear();
    for (size_t k = 0; k < all.size(); k++) {
        if (k == 0 || all[k].first != all[k - 1].first) {
            if (k) P.cptr.push_back((int)P.con.size());
            P.cell.push_back(all[k].first);
        }
        P.con.push_back(all[k].second);
    }
    P.cptr.push_back((int)P.con.size());
    const Geo &G = R.G;
    P.colptr.assign((size_t)G.NX * G.NY + 1, 0);
    for (size_t q : P.cell) P.colptr[(q / G.plane) * G.NX + (q % G.plane) / G.LDZ + 1]++;
    for (size_t l = 1; l < P.colptr.size(); l++) P.colptr[l] += P.colptr[l - 1];
}

static void check_shot3(const sfd_shot3d &s) {
    if (s.nt < 1 || s.ntr < 1 || !s.wavelet || !s.rz || !s.rx || !s.ry)
        throw EngineError(SFD_EINVAL, "3D shot needs nt >= 1, >= 1 receiver and its arrays");
}

// The shot's window and padded geometry, the window -> model maps and the
// stencil coefficients (no device work).
static void geom3(sfd_ctx3d &h, const sfd_shot3d &s, Run3 &R) {
    const sfd_grid3d &g = h.g;
    R.g = &g;
    window3(g, s, &R.oxw, &R.nxw, &R.oyw, &R.nyw);
    if (R.nxw < 1 || R.nyw < 1) throw EngineError(SFD_EINVAL, "empty shot window");
    Geo &G = R.G;
    G.NZ = g.nz + 2 * g.ntaper + 8;
    G.NX = R.nxw + 2 * g.ntaper + 8;
    G.NY = R.nyw + 2 * g.ntaper + 8;
    G.LDZ = (G.NZ + 31) / 32 * 32;
    G.plane = (size_t)G.LDZ * G.NX;
    R.nf = G.plane * G.NY;
    R.i0 = 4 + g.ntaper;   // padded index of window cell 1 on every axis
    R.fz = G.NZ - 8; R.fx = G.NX - 8; R.fy = G.NY - 8;
    R.rho = !h.rho_h.empty();
    // window columns / lines -> model columns / lines (mod_window_pad, nearest, clamped)
    R.colsrc.assign(R.nxw, 0);
    R.linesrc.assign(R.nyw, 0);
    for (int j = 1; j <= R.nxw; j++) {
        const int ix = nint3(((float)(j - 1) * g.dx + R.oxw - g.ox) / g.dx) + 1;
        R.colsrc[j - 1] = std::min(std::max(ix, 1), g.nx) - 1;
    }
    for (int k = 1; k <= R.nyw; k++) {
        const int iy = nint3(((float)(k - 1) * g.dy + R.oyw - g.oy) / g.dy) + 1;
        R.linesrc[k - 1] = std::min(std::max(iy, 1), g.ny) - 1;
    }
    R.dzi = 1.f / g.dz;
    R.dxi = 1.f / g.dx;
    R.dyi = 1.f / g.dy;
    if (R.rho) {   // FDcoefs.f90:12-23 over delta (coefpower 1)
        const float r[5] = {0.f, 1225.f / 1024.f, -(1225.f / (1024.f * 15.f)), 1225.f / (1024.f * 125.f),
                            -(1225.f / (1024.f * 1715.f))};
        for (int k = 0; k < 5; k++) {
            R.C.c[k] = r[k] * R.dzi;
            R.C.c[5 + k] = r[k] * R.dxi;
            R.C.c[10 + k] = r[k] * R.dyi;
        }
    } else {   // FDcoefs.f90:25-36 over delta^2
        const float cc[5] = {-205.f / 72.f, 8.f / 5, -1.f / 5.f, 8.f / 315.f, -1.f / 560.f};
        for (int k = 0; k < 5; k++) {
            R.C.c[k] = cc[k] * (R.dzi * R.dzi);
            R.C.c[5 + k] = cc[k] * (R.dxi * R.dxi);
            R.C.c[10 + k] = cc[k] * (R.dyi * R.dyi);
        }
    }
}

// Lay out the shot's device memory: five levels, models, Higdon tables, the
// window maps, the density arrays, then `extra` bytes at *extra_off (the
// caller's tables and outputs).
static void alloc3(sfd_ctx3d &h, Run3 &R, size_t extra, size_t *extra_off) {
    Layout L;
    size_t ou[5];
    for (int k = 0; k < 5; k++) ou[k] = L.take(R.nf * 4);
    const size_t ovel = L.take(R.nf * 4), ov2 = L.take(R.nf * 4);
    const size_t ogz = L.take((size_t)16 * R.fx * R.fy * 4), ogx = L.take((size_t)16 * R.fz * R.fy * 4),
                 ogy = L.take((size_t)16 * R.fz * R.fx * 4);
    const size_t ocol = L.take(R.nxw * 4), olin = L.take(R.nyw * 4);
    size_t orho = 0, odelp = 0, otmp = 0, osg[3] = {0, 0, 0};
    if (R.rho) {
        orho = L.take(R.nf * 4);
        odelp = L.take(R.nf * 4);
        otmp = L.take(R.nf * 4);
        for (int k = 0; k < 3; k++) osg[k] = L.take(R.nf * 4);
    }
    *extra_off = L.take(extra);
    R.base = h.reserve(L.off);
    for (int k = 0; k < 5; k++) R.u[k] = (float *)(R.base + ou[k]);
    R.vel = (float *)(R.base + ovel);
    R.v2 = (float *)(R.base + ov2);
    R.gz = (float *)(R.base + ogz);
    R.gx = (float *)(R.base + ogx);
    R.gy = (float *)(R.base + ogy);
    R.dcol = (const int *)(R.base + ocol);
    R.dlin = (const int *)(R.base + olin);
    HIPCHK(hipMemcpyAsync(R.base + ocol, R.colsrc.data(), R.nxw * 4, hipMemcpyHostToDevice, h.st));
    HIPCHK(hipMemcpyAsync(R.base + olin, R.linesrc.data(), R.nyw * 4, hipMemcpyHostToDevice, h.st));
    if (R.rho) {
        R.rhop = (float *)(R.base + orho);
        R.delp = (float *)(R.base + odelp);
        R.tmp = (float *)(R.base + otmp);
        for (int k = 0; k < 3; k++) R.sg[k] = (float *)(R.base + osg[k]);
    }
}

// The pass's model: padded vel and v^2, Higdon tables (Propagator.f90:152-154
// computes them at every call from the model it is given).
static void models3(sfd_ctx3d &h, Run3 &R, const float *d_model) {
    const sfd_grid3d &g = h.g;
    const Geo &G = R.G;
    const dim3 pg((G.NZ + 255) / 256, G.NX, G.NY);
    hipLaunchKernelGGL(pad3_kernel, pg, dim3(256), 0, h.st, d_model, g.nz, g.nx, R.dcol, R.nxw, R.dlin, R.nyw,
                       g.ntaper, G, R.vel, R.v2);
    const int mx = std::max(R.fx, R.fz);
    hipLaunchKernelGGL(higdon3_kernel, dim3((mx + 255) / 256, std::max(R.fy, R.fx), 3), dim3(256), 0, h.st,
                       (const float *)R.vel, G, g.dt, g.dz, g.dx, g.dy, R.gz, R.gx, R.gy);
    HIPCHK(hipGetLastError());
}

// A new propagator call: zeroed levels (and gradient arrays); with a density
// the padded rho and 1/rho2 (Interpolate: z pass rho -> delp, x pass delp ->
// tmp, y pass tmp -> delp, stored as 1/rho2).
static void start_pass3(sfd_ctx3d &h, Run3 &R, bool density_tables) {
    for (int k = 0; k < 5; k++) HIPCHK(hipMemsetAsync(R.u[k], 0, R.nf * 4, h.st));
    if (!R.rho) return;
    for (int k = 0; k < 3; k++) HIPCHK(hipMemsetAsync(R.sg[k], 0, R.nf * 4, h.st));
    if (!density_tables) return;
    const sfd_grid3d &g = h.g;
    const Geo &G = R.G;
    const dim3 pg((G.NZ + 255) / 256, G.NX, G.NY);
    hipLaunchKernelGGL(pad3_kernel, pg, dim3(256), 0, h.st, (const float *)h.d_rho, g.nz, g.nx, R.dcol, R.nxw, R.dlin,
                       R.nyw, g.ntaper, G, R.rhop, (float *)nullptr);
    Sinc9 S9;
    sfd_host_mksinc(S9.w, 9, 0.5f);
    hipLaunchKernelGGL(rho_half3_kernel, pg, dim3(256), 0, h.st, (const float *)R.rhop, R.delp, G, S9, 0, 0);
    hipLaunchKernelGGL(rho_half3_kernel, pg, dim3(256), 0, h.st, (const float *)R.delp, R.tmp, G, S9, 1, 0);
    hipLaunchKernelGGL(rho_half3_kernel, pg, dim3(256), 0, h.st, (const float *)R.tmp, R.delp, G, S9, 2, 1);
    HIPCHK(hipGetLastError());
}

// Device copy of an injection plan in `mem` (laid out by the caller).
struct InjDev {
    const size_t *cell = nullptr;
    const int *cptr = nullptr, *colptr = nullptr;
    Box3 box{1, 0, 1, 0, 1, 0};
    const InjC *con = nullptr;
    const float *v2r = nullptr, *traces = nullptr;
    int ncell = 0, nt = 0;
};
static size_t inj_bytes(const InjPlan &P, size_t ntrace_samples) {
    Layout L;
    L.take(P.cell.size() * sizeof(size_t));
    L.take(P.cptr.size() * 4);
    L.take(P.con.size() * sizeof(InjC));
    L.take(P.v2r.size() * 4);
    L.take(ntrace_samples * 4);
    L.take(P.colptr.size() * 4);
    return L.off + 256;
}
static InjDev upload_inj(sfd_ctx3d &h, char *mem, const InjPlan &P, const float *traces, int ntraces, int nt,
                         const Geo &G) {
    Layout L;
    const size_t o0 = L.take(P.cell.size() * sizeof(size_t)), o1 = L.take(P.cptr.size() * 4),
                 o2 = L.take(P.con.size() * sizeof(InjC)), o3 = L.take(P.v2r.size() * 4),
                 o4 = L.take((size_t)ntraces * nt * 4), o5 = L.take(P.colptr.size() * 4);
    HIPCHK(hipMemcpyAsync(mem + o0, P.cell.data(), P.cell.size() * sizeof(size_t), hipMemcpyHostToDevice, h.st));
    HIPCHK(hipMemcpyAsync(mem + o1, P.cptr.data(), P.cptr.size() * 4, hipMemcpyHostToDevice, h.st));
    HIPCHK(hipMemcpyAsync(mem + o2, P.con.data(), P.con.size() * sizeof(InjC), hipMemcpyHostToDevice, h.st));
    HIPCHK(hipMemcpyAsync(mem + o3, P.v2r.data(), P.v2r.size() * 4, hipMemcpyHostToDevice, h.st));
    HIPCHK(hipMemcpyAsync(mem + o4, traces, (size_t)ntraces * nt * 4, hipMemcpyHostToDevice, h.st));
    HIPCHK(hipMemcpyAsync(mem + o5, P.colptr.data(), P.colptr.size() * 4, hipMemcpyHostToDevice, h.st));
    InjDev D;
    D.cell = (const size_t *)(mem + o0);
    D.cptr = (const int *)(mem + o1);
    D.con = (const InjC *)(mem + o2);
    D.v2r = (const float *)(mem + o3);
    D.traces = (const float *)(mem + o4);
    D.colptr = P.cell.empty() ? nullptr : (const int *)(mem + o5);
    if (!P.cell.empty()) {
        D.box = Box3{1 << 30, -1, 1 << 30, -1, 1 << 30, -1};
        for (size_t q : P.cell) {
            const int y = (int)(q / G.plane), x = (int)((q % G.plane) / G.LDZ), z = (int)(q % G.LDZ);
            D.box = Box3{std::min(D.box.z0, z), std::max(D.box.z1, z), std::min(D.box.x0, x),
                         std::max(D.box.x1, x), std::min(D.box.y0, y), std::max(D.box.y1, y)};
        }
    }
    D.ncell = (int)P.cell.size();
    D.nt = nt;
    return D;
}

// Stencil, injection, time update and the Higdon bands of one step, then the
// level rotation (Propagator.f90:199-243).
constexpr int kFusedPlanes = 4;   // y planes per workgroup of step3_fused_kernel
static void step3(sfd_ctx3d &h, Run3 &R, float **buf, const InjDev &I, int it) {
    const Geo &G = R.G;
    const dim3 sb(256);
    const float c12dt2 = 12.f * (h.g.dt * h.g.dt), div11 = 1.f / 11.f;
    if (R.rho) {
        hipLaunchKernelGGL(rho_grad3_kernel, dim3((G.NZ + 63) / 64, (G.NX + 3) / 4, G.NY), sb, 0, h.st,
                           (const float *)buf[3], (const float *)R.delp, R.sg[0], R.sg[1], R.sg[2], G, R.C, R.dzi,
                           R.dxi, R.dyi);
        const dim3 fg((R.fz + F3_TZ - 1) / F3_TZ, (R.fx + 7) / 8, (R.fy + kFusedPlanes - 1) / kFusedPlanes);
        hipLaunchKernelGGL(step3_rho_fused_kernel<2>, fg, sb, 0, h.st, (const float *)buf[0], (const float *)buf[1],
                           (const float *)buf[2], (const float *)buf[3], buf[4], (const float *)R.sg[0],
                           (const float *)R.sg[1], (const float *)R.sg[2], (const float *)R.v2,
                           (const float *)R.rhop, G, R.C, c12dt2, div11, kFusedPlanes,
                           I.ncell > 0 ? I.box : Box3{1, 0, 1, 0, 1, 0}, I.colptr, I.cell, I.cptr, I.con, I.v2r,
                           I.traces, I.nt, it);
    } else {
        // 2 columns per lane (5 waves per SIMD) and 4 planes per workgroup
        // measured fastest on a 340^3 FD region (337 us per step against 347
        // for 8 planes and 349 for 4 columns per lane)
        const int yb = kFusedPlanes;
        const dim3 fg((R.fz + F3_TZ - 1) / F3_TZ, (R.fx + 7) / 8, (R.fy + yb - 1) / yb);
        hipLaunchKernelGGL((step3_fused_kernel<2, false, false>), fg, sb, 0, h.st, (const float *)buf[0],
                           (const float *)buf[1], (const float *)buf[2], (const float *)buf[3], buf[4],
                           (const float *)R.v2, G, R.C, c12dt2, div11, yb,
                           I.ncell > 0 ? I.box : Box3{1, 0, 1, 0, 1, 0}, I.colptr, I.cell, I.cptr, I.con, I.v2r,
                           I.traces, I.nt, it, Born3{});
    }
    hipLaunchKernelGGL(zband3s_kernel, dim3((R.fx + 7) / 8, R.fy), dim3(256), 0, h.st, buf[4], (const float *)buf[3],
                       (const float *)buf[2], (const float *)R.gz, G);
    hipLaunchKernelGGL(xband3_kernel, dim3((R.fz + 63) / 64, R.fy), dim3(64), 0, h.st, buf[4], (const float *)buf[3],
                       (const float *)buf[2], (const float *)R.gx, G);
    hipLaunchKernelGGL(yband3_kernel, dim3((R.fz + 63) / 64, R.fx), dim3(64), 0, h.st, buf[4], (const float *)buf[3],
                       (const float *)buf[2], (const float *)R.gy, G);
    float *t = buf[0];
    buf[0] = buf[1]; buf[1] = buf[2]; buf[2] = buf[3]; buf[3] = buf[4]; buf[4] = t;
}

// readsou with withRho (ReadSouVelRho.f90:328-347): the source shifted half a time step.
static std::vector<float> source3(const sfd_ctx3d &h, const sfd_shot3d &s) {
    std::vector<float> w(s.wavelet, s.wavelet + s.nt);
    if (h.rho_h.empty()) return w;
    float sinc[64];
    const int hl = h.g.lsinc / 2;
    sfd_host_mksinc(sinc, h.g.lsinc, 0.5f);
    for (int it = 1; it <= s.nt; it++) {
        float fscale = 1.f, fdum = 0.f;
        for (int l = -hl; l <= hl; l++) {
            if (it + l >= 1 && it + l <= s.nt) fdum = fdum + sinc[hl + l] * s.wavelet[it + l - 1];
            else fscale = fscale - sinc[hl + l];
        }
        w[it - 1] = fdum / fscale;
    }
    return w;
}

static void forward3(sfd_ctx3d &h, const sfd_shot3d &s, float *traces_out, int kind) {
    const sfd_grid3d &g = h.g;
    check_shot3(s);
    Run3 R;
    geom3(h, s, R);
    InjPlan ip;
    plan_inj3(R, h.vel_h, h.rho_h, 1, &s.sz, &s.sx, &s.sy, ip);
    // receivers (Extraction.f90:71-162,217-246 -> :276-337)
    std::vector<Rec3> rec(s.ntr);
    const size_t trace_bytes = (size_t)s.nt * s.ntr * 4;
    const size_t rec_bytes = sizeof(Rec3) * s.ntr;
    for (int r = 0; r < s.ntr; r++) {
        int ic[3];
        float dc[3];
        find_i_d3(&ic[0], &dc[0], s.rz[r], g.oz, g.dz);
        find_i_d3(&ic[1], &dc[1], s.rx[r], R.oxw, g.dx);
        find_i_d3(&ic[2], &dc[2], s.ry[r], R.oyw, g.dy);
        Rec3 &e = rec[r];
        e.iz = ic[0] - 1 + R.i0;
        e.ix = ic[1] - 1 + R.i0;
        e.iy = ic[2] - 1 + R.i0;
        if (e.iz - 3 < 0 || e.iz + 3 >= R.G.NZ || e.ix - 3 < 0 || e.ix + 3 >= R.G.NX || e.iy - 3 < 0 ||
            e.iy + 3 >= R.G.NY)
            throw EngineError(SFD_EINVAL, "receiver outside the padded grid");
        e.gain = kind == SFD_EXTRACT_SINC_V2 ? R.at(h.vel_h, ic[0], ic[1], ic[2])
                 : kind == SFD_EXTRACT_SINC_LSRTM ? 1.f / (g.dz * g.dx * g.dy)
                                                  : 1.f;
        sfd_host_mksinc(e.wz, 7, dc[0] / g.dz);
        sfd_host_mksinc(e.wx, 7, dc[1] / g.dx);
    }
    const size_t extra = trace_bytes + 256 + rec_bytes + 256 + inj_bytes(ip, s.nt);
    size_t extra_off = 0;
    alloc3(h, R, extra, &extra_off);
    char *ex = R.base + extra_off;
    float *trc = (float *)ex;
    Rec3 *drec = (Rec3 *)(ex + ((trace_bytes + 255) & ~(size_t)255));
    char *injmem = (char *)drec + ((rec_bytes + 255) & ~(size_t)255);
    HIPCHK(hipMemsetAsync(trc, 0, trace_bytes, h.st));
    HIPCHK(hipMemcpyAsync(drec, rec.data(), rec_bytes, hipMemcpyHostToDevice, h.st));
    const std::vector<float> wav = source3(h, s);
    const InjDev I = upload_inj(h, injmem, ip, wav.data(), 1, s.nt, R.G);
    models3(h, R, h.d_vel);
    start_pass3(h, R, true);
    const int nearest = kind == SFD_EXTRACT_NEAREST;
    float *buf[5] = {R.u[0], R.u[1], R.u[2], R.u[3], R.u[4]};   // u_-1, u0, u1, u2, u3 (FDswaptimes.f90:14-24)
    for (int it = 1; it <= s.nt; it++) {
        extract3(h.st, R.G, buf[3], (const Rec3 *)drec, s.ntr, nearest, trc, s.nt, it);
        step3(h, R, buf, I, it);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(traces_out, trc, trace_bytes, hipMemcpyDeviceToHost, h.st));
    HIPCHK(hipStreamSynchronize(h.st));
}

// RTM_to_memory with genpar%twoD false (to_disk_to_memory_RTM.f90:358-492).
static void rtm3(sfd_ctx3d &h, const sfd_shot3d &s, const float *traces_in, float *image, float *illum) {
    const sfd_grid3d &g = h.g;
    check_shot3(s);
    if (g.snapi < 1) throw EngineError(SFD_EINVAL, "snapi must be >= 1");
    Run3 R;
    geom3(h, s, R);
    InjPlan src, recp;
    plan_inj3(R, h.vel_h, h.rho_h, 1, &s.sz, &s.sx, &s.sy, src);
    plan_inj3(R, h.velb_h.empty() ? h.vel_h : h.velb_h, h.rho_h, s.ntr, s.rz, s.rx, s.ry, recp);
    const int ntsnap = s.nt / g.snapi;
    const size_t nwin = (size_t)g.nz * R.nxw * R.nyw;
    Layout E;
    const size_t o_snap = E.take(nwin * std::max(ntsnap, 1) * 4), o_img = E.take(nwin * 4), o_ill = E.take(nwin * 4),
                 o_tx = E.take(R.nxw * 4), o_ty = E.take(R.nyw * 4), o_src = E.take(inj_bytes(src, s.nt)),
                 o_rec = E.take(inj_bytes(recp, (size_t)s.ntr * s.nt));
    size_t extra_off = 0;
    alloc3(h, R, E.off, &extra_off);
    char *ex = R.base + extra_off;
    float *snaps = (float *)(ex + o_snap), *img = (float *)(ex + o_img), *ill = (float *)(ex + o_ill);
    HIPCHK(hipMemsetAsync(img, 0, nwin * 4, h.st));
    HIPCHK(hipMemsetAsync(ill, 0, nwin * 4, h.st));
    std::vector<float> tx(R.nxw), ty(R.nyw, 1.f);
    sfd_host_taper(tx.data(), R.nxw);
    if (R.nyw > 1) sfd_host_taper(ty.data(), R.nyw);
    HIPCHK(hipMemcpyAsync(ex + o_tx, tx.data(), R.nxw * 4, hipMemcpyHostToDevice, h.st));
    HIPCHK(hipMemcpyAsync(ex + o_ty, ty.data(), R.nyw * 4, hipMemcpyHostToDevice, h.st));
    const std::vector<float> wav = source3(h, s);
    const InjDev Is = upload_inj(h, ex + o_src, src, wav.data(), 1, s.nt, R.G);
    const InjDev Ir = upload_inj(h, ex + o_rec, recp, traces_in, s.ntr, s.nt, R.G);
    const dim3 wg((g.nz + 255) / 256, R.nxw, R.nyw);
    // forward: snapshots every snapi steps
    models3(h, R, h.d_vel);
    start_pass3(h, R, true);
    float *buf[5] = {R.u[0], R.u[1], R.u[2], R.u[3], R.u[4]};
    int counter = 0;
    for (int it = 1; it <= s.nt; it++) {
        if (it % g.snapi == 0)
            hipLaunchKernelGGL(snap3_kernel, wg, dim3(256), 0, h.st, (const float *)buf[3], R.G, g.nz, R.nxw, R.i0,
                               snaps + (size_t)(counter++) * nwin);
        step3(h, R, buf, Is, it);
    }
    // backward: a new propagator call on the backward model, receivers as sources
    if (h.d_velb) models3(h, R, h.d_velb);
    start_pass3(h, R, false);
    for (int k = 0; k < 5; k++) buf[k] = R.u[k];
    counter = 0;
    for (int it = s.nt; it >= 1; it--) {
        if (it % g.snapi == 0) {
            counter++;
            const int index = ntsnap - counter + 1;
            hipLaunchKernelGGL(image3_kernel, wg, dim3(256), 0, h.st, (const float *)buf[3], R.G, g.nz, R.nxw, R.i0,
                               (const float *)(snaps + (size_t)(index - 1) * nwin), (const float *)(ex + o_tx),
                               (const float *)(ex + o_ty), img, ill);
        }
        step3(h, R, buf, Ir, it);
    }
    HIPCHK(hipGetLastError());
    // mod_copy_image (ExtractPadModel.f90:662-695): a zeroed model-size sum of the window, then added
    std::vector<float> wimg(nwin), will(nwin);
    HIPCHK(hipMemcpyAsync(wimg.data(), img, nwin * 4, hipMemcpyDeviceToHost, h.st));
    HIPCHK(hipMemcpyAsync(will.data(), ill, nwin * 4, hipMemcpyDeviceToHost, h.st));
    HIPCHK(hipStreamSynchronize(h.st));
    const size_t N = (size_t)g.nz * g.nx * g.ny;
    for (int pass = 0; pass < 2; pass++) {
        float *dst = pass == 0 ? image : illum;
        const std::vector<float> &w = pass == 0 ? wimg : will;
        if (!dst) continue;
        std::vector<float> tmp(N, 0.f);
        for (int k = 1; k <= R.nyw; k++) {
            const int iy = nint3(((float)(k - 1) * g.dy + R.oyw - g.oy) / g.dy) + 1;
            if (iy < 1 || iy > g.ny) continue;
            for (int j = 1; j <= R.nxw; j++) {
                const int ix = nint3(((float)(j - 1) * g.dx + R.oxw - g.ox) / g.dx) + 1;
                if (ix < 1 || ix > g.nx) continue;
                float *d = tmp.data() + (size_t)(ix - 1) * g.nz + (size_t)(iy - 1) * g.nz * g.nx;
                const float *ww = w.data() + (size_t)(j - 1) * g.nz + (size_t)(k - 1) * g.nz * R.nxw;
                for (int i = 0; i < g.nz; i++) d[i] = d[i] + ww[i];
            }
        }
        for (size_t q = 0; q < N; q++) dst[q] = dst[q] + tmp[q];
    }
}

// Compute_OF_RES_3D (Objective_Function_Residual_Adjsource.f90:42-67) on one
// shot's traces, in trace order, with fct_compute / gdt_compute
// (libs/libinv/Norm.f90:12-141): r = w (dobs - dmod); misfit += the trace's
// norm (L2: FP64 products over 2; the others FP32 terms summed in FP32 in
// time order); adjoint source w r, through gdt for the non-L2 norms.  One
// THRESH for the shot: data_threshold, or when 0 (and the norm uses it) the
// max |r| of the first trace whose residual is not all zero.
// f: the running misfit over every trace of the call, added to in trace order.
static void residual3(int norm, float thresh0, int nt, int ntr, const float *dobs, const float *mute, float *d,
                      double &f) {
    const bool uses = norm == SFD_NORM_HUBER || norm == SFD_NORM_CAUCHY || norm == SFD_NORM_HYPERBOLIC;
    float t = thresh0;
    for (int r = 0; r < ntr; r++) {
        float *x = d + (size_t)r * nt;
        const float *ob = dobs + (size_t)r * nt, *mu = mute ? mute + (size_t)r * nt : nullptr;
        float m = 0.f;
        for (int it = 0; it < nt; it++) {
            x[it] = (mu ? mu[it] : 1.0f) * (ob[it] - x[it]);
            m = std::max(m, fabsf(x[it]));
        }
        if (uses && t == 0.f) t = m;
        if (norm == SFD_NORM_L2) {
            double a = 0.0;
            for (int it = 0; it < nt; it++) a = a + (double)x[it] * (double)x[it];
            f = f + a / 2;
        } else {
            float acc = 0.f;
            for (int it = 0; it < nt; it++) {
                const float v = x[it], av = fabsf(v);
                float term;
                if (norm == SFD_NORM_L1) term = av;
                else if (norm == SFD_NORM_HUBER) term = av > t ? av - t / 2.0f : v * v / (2.0f * t);
                else if (norm == SFD_NORM_CAUCHY) { const float qq = v / t; term = logf(1.0f + qq * qq) / 2.0f; }
                else term = sqrtf(t * t + v * v) - t;
                acc = acc + term;
            }
            f = f + (double)acc;
        }
        for (int it = 0; it < nt; it++) {
            float y = (mu ? mu[it] : 1.0f) * x[it];
            if (norm == SFD_NORM_L1) y = y < 0.f ? -1.f : (y > 0.f ? 1.f : 0.f);
            else if (norm == SFD_NORM_HUBER) { const float qq = y / t; y = qq > 1.f ? 1.f : (qq < -1.f ? -1.f : qq); }
            else if (norm == SFD_NORM_CAUCHY) y = y / (t * t + y * y);
            else if (norm == SFD_NORM_HYPERBOLIC) y = (y / t) / sqrtf(1.0f + y * y / (t * t));
            x[it] = y;
        }
    }
}

// Adds a window field (nz, nxw, nyw) into a model field (nz, nx, ny) at the
// nearest model columns and lines (mod_copy_gradient_to_disk / mod_copy_image).
static void copy_window3(const sfd_grid3d &g, const Run3 &R, const float *w, float *dst) {
    for (int k = 1; k <= R.nyw; k++) {
        const int iy = nint3(((float)(k - 1) * g.dy + R.oyw - g.oy) / g.dy) + 1;
        if (iy < 1 || iy > g.ny) continue;
        for (int j = 1; j <= R.nxw; j++) {
            const int ix = nint3(((float)(j - 1) * g.dx + R.oxw - g.ox) / g.dx) + 1;
            if (ix < 1 || ix > g.nx) continue;
            float *d = dst + (size_t)(ix - 1) * g.nz + (size_t)(iy - 1) * g.nz * g.nx;
            const float *ww = w + (size_t)(j - 1) * g.nz + (size_t)(k - 1) * g.nz * R.nxw;
            for (int i = 0; i < g.nz; i++) d[i] = d[i] + ww[i];
        }
    }
}

// AFWI_to_memory with genpar%twoD false (to_disk_to_memory_AFWI.f90:229-415),
// one shot: forward with gain-1 sinc extraction and snapshots, the residual
// on the host, reverse time with the residual injected and the AFWI or v-rho
// imaging; the windows are added into grad (vp, then rho) and illum.
static void afwi3(sfd_ctx3d &h, const sfd_shot3d &s, const float *dobs, const float *mute, int norm, float thresh,
                  float *grad, float *illum, double &f) {
    const sfd_grid3d &g = h.g;
    check_shot3(s);
    Run3 R;
    geom3(h, s, R);
    if (R.nxw < 2) throw EngineError(SFD_EINVAL, "the imaging's one-sided x difference needs two window columns");
    InjPlan src, recp;
    plan_inj3(R, h.vel_h, h.rho_h, 1, &s.sz, &s.sx, &s.sy, src);
    plan_inj3(R, h.vel_h, h.rho_h, s.ntr, s.rz, s.rx, s.ry, recp);
    std::vector<Rec3> rec(s.ntr);
    for (int r = 0; r < s.ntr; r++) {   // Extraction_array_sinc_afwi: gain 1
        int ic[3];
        float dc[3];
        find_i_d3(&ic[0], &dc[0], s.rz[r], g.oz, g.dz);
        find_i_d3(&ic[1], &dc[1], s.rx[r], R.oxw, g.dx);
        find_i_d3(&ic[2], &dc[2], s.ry[r], R.oyw, g.dy);
        Rec3 &e = rec[r];
        e.iz = ic[0] - 1 + R.i0;
        e.ix = ic[1] - 1 + R.i0;
        e.iy = ic[2] - 1 + R.i0;
        e.gain = 1.f;
        sfd_host_mksinc(e.wz, 7, dc[0] / g.dz);
        sfd_host_mksinc(e.wx, 7, dc[1] / g.dx);
    }
    const int ntsnap = s.nt / g.snapi;
    const size_t nwin = (size_t)g.nz * R.nxw * R.nyw, ntr_nt = (size_t)s.nt * s.ntr;
    Layout E;
    const size_t o_snap = E.take(nwin * std::max(ntsnap, 1) * 4), o_img = E.take(nwin * 4),
                 o_img2 = E.take(nwin * 4), o_ill = E.take(nwin * 4), o_tx = E.take(R.nxw * 4),
                 o_ty = E.take(R.nyw * 4), o_trc = E.take(ntr_nt * 4), o_rec = E.take(sizeof(Rec3) * s.ntr),
                 o_src = E.take(inj_bytes(src, s.nt)), o_rinj = E.take(inj_bytes(recp, ntr_nt));
    size_t extra_off = 0;
    alloc3(h, R, E.off, &extra_off);
    char *ex = R.base + extra_off;
    float *snaps = (float *)(ex + o_snap), *img = (float *)(ex + o_img), *img2 = (float *)(ex + o_img2),
          *ill = (float *)(ex + o_ill), *trc = (float *)(ex + o_trc);
    for (size_t o : {o_img, o_img2, o_ill}) HIPCHK(hipMemsetAsync(ex + o, 0, nwin * 4, h.st));
    HIPCHK(hipMemsetAsync(trc, 0, ntr_nt * 4, h.st));
    std::vector<float> tx(R.nxw), ty(R.nyw, 1.f);
    sfd_host_taper(tx.data(), R.nxw);
    if (R.nyw > 1) sfd_host_taper(ty.data(), R.nyw);
    HIPCHK(hipMemcpyAsync(ex + o_tx, tx.data(), R.nxw * 4, hipMemcpyHostToDevice, h.st));
    HIPCHK(hipMemcpyAsync(ex + o_ty, ty.data(), R.nyw * 4, hipMemcpyHostToDevice, h.st));
    HIPCHK(hipMemcpyAsync(ex + o_rec, rec.data(), sizeof(Rec3) * s.ntr, hipMemcpyHostToDevice, h.st));
    const std::vector<float> wav = source3(h, s);
    const InjDev Is = upload_inj(h, ex + o_src, src, wav.data(), 1, s.nt, R.G);
    const dim3 wg((g.nz + 255) / 256, R.nxw, R.nyw);
    models3(h, R, h.d_vel);
    start_pass3(h, R, true);
    float *buf[5] = {R.u[0], R.u[1], R.u[2], R.u[3], R.u[4]};
    int counter = 0;
    for (int it = 1; it <= s.nt; it++) {
        extract3(h.st, R.G, buf[3], (const Rec3 *)(ex + o_rec), s.ntr, 0, trc, s.nt, it);
        if (it % g.snapi == 0)
            hipLaunchKernelGGL(snap3_kernel, wg, dim3(256), 0, h.st, (const float *)buf[3], R.G, g.nz, R.nxw, R.i0,
                               snaps + (size_t)(counter++) * nwin);
        step3(h, R, buf, Is, it);
    }
    HIPCHK(hipGetLastError());
    std::vector<float> d(ntr_nt);
    HIPCHK(hipMemcpyAsync(d.data(), trc, ntr_nt * 4, hipMemcpyDeviceToHost, h.st));
    HIPCHK(hipStreamSynchronize(h.st));
    residual3(norm, thresh, s.nt, s.ntr, dobs, mute, d.data(), f);
    const InjDev Ir = upload_inj(h, ex + o_rinj, recp, d.data(), s.ntr, s.nt, R.G);
    start_pass3(h, R, false);
    for (int k = 0; k < 5; k++) buf[k] = R.u[k];
    const float dt2 = (g.dt * (float)g.snapi) * (g.dt * (float)g.snapi);
    counter = 0;
    for (int it = s.nt; it >= 1; it--) {
        if (it % g.snapi == 0) {
            counter++;
            const int index = ntsnap - counter + 1;
            const int im1 = index - 1 > 1 ? index - 1 : 1, ip1 = index + 1 < ntsnap ? index + 1 : ntsnap;
            const float *S = snaps + (size_t)(index - 1) * nwin, *Sm = snaps + (size_t)(im1 - 1) * nwin,
                        *Sp = snaps + (size_t)(ip1 - 1) * nwin;
            if (R.rho)
                hipLaunchKernelGGL(rhovp3_image_kernel, wg, dim3(256), 0, h.st, (const float *)buf[3],
                                   (const float *)R.vel, (const float *)R.rhop, R.G, g.nz, R.nxw, R.nyw, R.i0, S, Sm,
                                   Sp, (const float *)(ex + o_tx), (const float *)(ex + o_ty), dt2, R.dzi, R.dxi,
                                   R.dyi, img, img2, ill);
            else
                hipLaunchKernelGGL(afwi3_image_kernel, wg, dim3(256), 0, h.st, (const float *)buf[3],
                                   (const float *)R.vel, R.G, g.nz, R.nxw, R.i0, S, Sm, Sp, (const float *)(ex + o_tx),
                                   (const float *)(ex + o_ty), dt2, img, ill);
        }
        step3(h, R, buf, Ir, it);
    }
    HIPCHK(hipGetLastError());
    const size_t N = (size_t)g.nz * g.nx * g.ny;
    std::vector<float> w(nwin);
    for (int pass = 0; pass < 3; pass++) {
        if (pass == 1 && !R.rho) continue;
        if (pass == 2 && !illum) continue;
        HIPCHK(hipMemcpyAsync(w.data(), pass == 0 ? img : pass == 1 ? img2 : ill, nwin * 4, hipMemcpyDeviceToHost,
                              h.st));
        HIPCHK(hipStreamSynchronize(h.st));
        copy_window3(g, R, w.data(), pass == 0 ? grad : pass == 1 ? grad + N : illum);
    }
}

// The forward pass of LSRTM and Born in 3D (readsou with Born / LSRTM,
// ReadSouVelRho.f90:318-326): the second-differenced source by Injection_sinc,
// the forward stencil, u2 on the window every snapi steps into snaps.
static void fwd_snaps3(sfd_ctx3d &h, Run3 &R, const sfd_shot3d &s, const InjPlan &src, char *srcmem, float *snaps) {
    const sfd_grid3d &g = h.g;
    const size_t nwin = (size_t)g.nz * R.nxw * R.nyw;
    std::vector<float> w(s.wavelet, s.wavelet + s.nt);
    second_difference(w, g.dt);
    const InjDev Is = upload_inj(h, srcmem, src, w.data(), 1, s.nt, R.G);
    const dim3 wg((g.nz + 255) / 256, R.nxw, R.nyw);
    models3(h, R, h.d_vel);
    start_pass3(h, R, false);
    float *buf[5] = {R.u[0], R.u[1], R.u[2], R.u[3], R.u[4]};
    int counter = 0;
    for (int it = 1; it <= s.nt; it++) {
        if (it % g.snapi == 0)
            hipLaunchKernelGGL(snap3_kernel, wg, dim3(256), 0, h.st, (const float *)buf[3], R.G, g.nz, R.nxw, R.i0,
                               snaps + (size_t)(counter++) * nwin);
        step3(h, R, buf, Is, it);
    }
    HIPCHK(hipStreamSynchronize(h.st));   // w is released on return
}

static void tapers3(sfd_ctx3d &h, const Run3 &R, char *dtx, char *dty) {
    std::vector<float> tx(R.nxw), ty(R.nyw, 1.f);
    sfd_host_taper(tx.data(), R.nxw);
    if (R.nyw > 1) sfd_host_taper(ty.data(), R.nyw);
    HIPCHK(hipMemcpyAsync(dtx, tx.data(), R.nxw * 4, hipMemcpyHostToDevice, h.st));
    HIPCHK(hipMemcpyAsync(dty, ty.data(), R.nyw * 4, hipMemcpyHostToDevice, h.st));
    HIPCHK(hipStreamSynchronize(h.st));
}

// LSRTM_to_memory with genpar%twoD false (to_disk_to_memory_RTM.f90:494-579).
static void lsrtm3(sfd_ctx3d &h, const sfd_shot3d &s, const float *traces_in, float *image, float *illum) {
    const sfd_grid3d &g = h.g;
    check_shot3(s);
    Run3 R;
    geom3(h, s, R);
    InjPlan src, recp;
    plan_inj3(R, h.vel_h, h.rho_h, 1, &s.sz, &s.sx, &s.sy, src);
    plan_inj3(R, h.vel_h, h.rho_h, s.ntr, s.rz, s.rx, s.ry, recp);
    const float wls = 1.f / (g.dz * g.dx * g.dy);   // Injection_LSRTM_sinc: 1/product(delta(1:3))
    for (float &v : recp.v2r) v = wls;
    const int ntsnap = s.nt / g.snapi;
    const size_t nwin = (size_t)g.nz * R.nxw * R.nyw;
    Layout E;
    const size_t o_snap = E.take(nwin * std::max(ntsnap, 1) * 4), o_img = E.take(nwin * 4), o_ill = E.take(nwin * 4),
                 o_tx = E.take(R.nxw * 4), o_ty = E.take(R.nyw * 4),
                 o_src = E.take(inj_bytes(src, s.nt)), o_rec = E.take(inj_bytes(recp, (size_t)s.ntr * s.nt));
    size_t extra_off = 0;
    alloc3(h, R, E.off, &extra_off);
    char *ex = R.base + extra_off;
    float *snaps = (float *)(ex + o_snap), *img = (float *)(ex + o_img), *ill = (float *)(ex + o_ill);
    HIPCHK(hipMemsetAsync(img, 0, nwin * 4, h.st));
    HIPCHK(hipMemsetAsync(ill, 0, nwin * 4, h.st));
    tapers3(h, R, ex + o_tx, ex + o_ty);
    fwd_snaps3(h, R, s, src, ex + o_src, snaps);
    const InjDev Ir = upload_inj(h, ex + o_rec, recp, traces_in, s.ntr, s.nt, R.G);
    if (h.d_velb) models3(h, R, h.d_velb);
    start_pass3(h, R, false);
    float *buf[5] = {R.u[0], R.u[1], R.u[2], R.u[3], R.u[4]};
    const dim3 wg((g.nz + 255) / 256, R.nxw, R.nyw), sb(256);
    const dim3 fg((R.fz + F3_TZ - 1) / F3_TZ, (R.fx + 7) / 8, (R.fy + kFusedPlanes - 1) / kFusedPlanes);
    const float c12dt2 = 12.f * (g.dt * g.dt), div11 = 1.f / 11.f;
    int counter = 0;
    for (int it = s.nt; it >= 1; it--) {
        if (it % g.snapi == 0) {
            counter++;
            const int index = ntsnap - counter + 1;
            hipLaunchKernelGGL(lsrtm3_image_kernel, wg, dim3(256), 0, h.st, (const float *)buf[3],
                               (const float *)R.vel, R.G, g.nz, R.nxw, R.i0,
                               (const float *)(snaps + (size_t)(index - 1) * nwin), (const float *)(ex + o_tx),
                               (const float *)(ex + o_ty), img, ill);
        }
        // the adjoint-style stencil (FDderivatives.f90:780-819), then injection, update, bands
        hipLaunchKernelGGL((step3_fused_kernel<2, true, false>), fg, sb, 0, h.st, (const float *)buf[0],
                           (const float *)buf[1], (const float *)buf[2], (const float *)buf[3], buf[4],
                           (const float *)R.v2, R.G, R.C, c12dt2, div11, kFusedPlanes,
                           Ir.ncell > 0 ? Ir.box : Box3{1, 0, 1, 0, 1, 0}, Ir.colptr, Ir.cell, Ir.cptr, Ir.con,
                           Ir.v2r, Ir.traces, Ir.nt, it, Born3{});
        hipLaunchKernelGGL(zband3s_kernel, dim3((R.fx + 7) / 8, R.fy), dim3(256), 0, h.st, buf[4],
                           (const float *)buf[3], (const float *)buf[2], (const float *)R.gz, R.G);
        hipLaunchKernelGGL(xband3_kernel, dim3((R.fz + 63) / 64, R.fy), dim3(64), 0, h.st, buf[4],
                           (const float *)buf[3], (const float *)buf[2], (const float *)R.gx, R.G);
        hipLaunchKernelGGL(yband3_kernel, dim3((R.fz + 63) / 64, R.fx), dim3(64), 0, h.st, buf[4],
                           (const float *)buf[3], (const float *)buf[2], (const float *)R.gy, R.G);
        float *t = buf[0];
        buf[0] = buf[1]; buf[1] = buf[2]; buf[2] = buf[3]; buf[3] = buf[4]; buf[4] = t;
    }
    HIPCHK(hipGetLastError());
    // mod_copy_image: a zeroed model-size sum of the window, then added
    const size_t N = (size_t)g.nz * g.nx * g.ny;
    std::vector<float> w(nwin), tmpm;
    for (int pass = 0; pass < 2; pass++) {
        float *dst = pass == 0 ? image : illum;
        if (!dst) continue;
        HIPCHK(hipMemcpyAsync(w.data(), pass == 0 ? img : ill, nwin * 4, hipMemcpyDeviceToHost, h.st));
        HIPCHK(hipStreamSynchronize(h.st));
        tmpm.assign(N, 0.f);
        copy_window3(g, R, w.data(), tmpm.data());
        for (size_t q = 0; q < N; q++) dst[q] = dst[q] + tmpm[q];
    }
}

// BORN_to_memory with genpar%twoD false (to_disk_to_memory_BORN.f90:123-220).
static void born3(sfd_ctx3d &h, const sfd_shot3d &s, const float *d_refl, float *traces_out) {
    const sfd_grid3d &g = h.g;
    check_shot3(s);
    Run3 R;
    geom3(h, s, R);
    InjPlan src;
    plan_inj3(R, h.vel_h, h.rho_h, 1, &s.sz, &s.sx, &s.sy, src);
    std::vector<Rec3> rec(s.ntr);
    for (int r = 0; r < s.ntr; r++) {   // Extraction_array_LSRTM_sinc: gain 1/(dz dx dy)
        int ic[3];
        float dc[3];
        find_i_d3(&ic[0], &dc[0], s.rz[r], g.oz, g.dz);
        find_i_d3(&ic[1], &dc[1], s.rx[r], R.oxw, g.dx);
        find_i_d3(&ic[2], &dc[2], s.ry[r], R.oyw, g.dy);
        Rec3 &e = rec[r];
        e.iz = ic[0] - 1 + R.i0;
        e.ix = ic[1] - 1 + R.i0;
        e.iy = ic[2] - 1 + R.i0;
        if (e.iz - 3 < 0 || e.iz + 3 >= R.G.NZ || e.ix - 3 < 0 || e.ix + 3 >= R.G.NX || e.iy - 3 < 0 ||
            e.iy + 3 >= R.G.NY)
            throw EngineError(SFD_EINVAL, "receiver outside the padded grid");
        e.gain = 1.f / (g.dz * g.dx * g.dy);
        sfd_host_mksinc(e.wz, 7, dc[0] / g.dz);
        sfd_host_mksinc(e.wx, 7, dc[1] / g.dx);
    }
    const int ntsnap = s.nt / g.snapi;
    const size_t nwin = (size_t)g.nz * R.nxw * R.nyw, ntr_nt = (size_t)s.nt * s.ntr;
    Layout E;
    const size_t o_snap = E.take(nwin * std::max(ntsnap, 1) * 4), o_m = E.take(R.G.plane * R.G.NY * 4),
                 o_tx = E.take(R.nxw * 4), o_ty = E.take(R.nyw * 4), o_trc = E.take(ntr_nt * 4),
                 o_rec = E.take(sizeof(Rec3) * s.ntr), o_src = E.take(inj_bytes(src, s.nt));
    size_t extra_off = 0;
    alloc3(h, R, E.off, &extra_off);
    char *ex = R.base + extra_off;
    float *snaps = (float *)(ex + o_snap), *m = (float *)(ex + o_m), *trc = (float *)(ex + o_trc);
    HIPCHK(hipMemsetAsync(trc, 0, ntr_nt * 4, h.st));
    HIPCHK(hipMemcpyAsync(ex + o_rec, rec.data(), sizeof(Rec3) * s.ntr, hipMemcpyHostToDevice, h.st));
    tapers3(h, R, ex + o_tx, ex + o_ty);
    fwd_snaps3(h, R, s, src, ex + o_src, snaps);
    // readref + mod_window_pad: the reflectivity windowed and padded like vel
    const Geo &G = R.G;
    hipLaunchKernelGGL(pad3_kernel, dim3((G.NZ + 255) / 256, G.NX, G.NY), dim3(256), 0, h.st, d_refl, g.nz, g.nx,
                       R.dcol, R.nxw, R.dlin, R.nyw, g.ntaper, G, m, (float *)nullptr);
    if (h.d_velb) models3(h, R, h.d_velb);
    start_pass3(h, R, false);
    float *buf[5] = {R.u[0], R.u[1], R.u[2], R.u[3], R.u[4]};
    const dim3 wg((g.nz + 255) / 256, R.nxw, R.nyw), sb(256);
    const dim3 fg((R.fz + F3_TZ - 1) / F3_TZ, (R.fx + 7) / 8, (R.fy + kFusedPlanes - 1) / kFusedPlanes);
    const float c12dt2 = 12.f * (g.dt * g.dt), div11 = 1.f / 11.f;
    int counter = 0;
    for (int it = 1; it <= s.nt; it++) {
        extract3(h.st, G, buf[3], (const Rec3 *)(ex + o_rec), s.ntr, 0, trc, s.nt, it);
        if (it % g.snapi == 0) {   // Injection_Born after the stencil, in the same pass
            const Born3 bo{(const float *)m, (const float *)R.vel, (const float *)(snaps + (size_t)(counter++) * nwin),
                           (const float *)(ex + o_tx), (const float *)(ex + o_ty), g.nz, R.nxw, R.nyw, R.i0};
            hipLaunchKernelGGL((step3_fused_kernel<2, false, true>), fg, sb, 0, h.st, (const float *)buf[0],
                               (const float *)buf[1], (const float *)buf[2], (const float *)buf[3], buf[4],
                               (const float *)R.v2, G, R.C, c12dt2, div11, kFusedPlanes, Box3{1, 0, 1, 0, 1, 0},
                               (const int *)nullptr, (const size_t *)nullptr, (const int *)nullptr,
                               (const InjC *)nullptr, (const float *)nullptr, (const float *)nullptr, 0, it, bo);
        } else {
            hipLaunchKernelGGL((step3_fused_kernel<2, false, false>), fg, sb, 0, h.st, (const float *)buf[0],
                               (const float *)buf[1], (const float *)buf[2], (const float *)buf[3], buf[4],
                               (const float *)R.v2, G, R.C, c12dt2, div11, kFusedPlanes, Box3{1, 0, 1, 0, 1, 0},
                               (const int *)nullptr, (const size_t *)nullptr, (const int *)nullptr,
                               (const InjC *)nullptr, (const float *)nullptr, (const float *)nullptr, 0, it, Born3{});
        }
        hipLaunchKernelGGL(zband3s_kernel, dim3((R.fx + 7) / 8, R.fy), dim3(256), 0, h.st, buf[4],
                           (const float *)buf[3], (const float *)buf[2], (const float *)R.gz, G);
        hipLaunchKernelGGL(xband3_kernel, dim3((R.fz + 63) / 64, R.fy), dim3(64), 0, h.st, buf[4],
                           (const float *)buf[3], (const float *)buf[2], (const float *)R.gx, G);
        hipLaunchKernelGGL(yband3_kernel, dim3((R.fz + 63) / 64, R.fx), dim3(64), 0, h.st, buf[4],
                           (const float *)buf[3], (const float *)buf[2], (const float *)R.gy, G);
        float *t = buf[0];
        buf[0] = buf[1]; buf[1] = buf[2]; buf[2] = buf[3]; buf[3] = buf[4]; buf[4] = t;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(traces_out, trc, ntr_nt * 4, hipMemcpyDeviceToHost, h.st));
    HIPCHK(hipStreamSynchronize(h.st));
}

static std::string g_create_err3d;   // message of the last failed sfd_create3d

extern "C" {

int sfd_create3d(sfd_handle3d *h, const sfd_grid3d *g) {
    if (!h || !g) return SFD_EINVAL;
    *h = nullptr;
    return abi_call(&g_create_err3d, [&]() {
        check3(*g);
        std::unique_ptr<sfd_ctx3d> c(new sfd_ctx3d());
        c->g = *g;
        HIPCHK(hipStreamCreate(&c->st));
        *h = c.release();
    });
}

int sfd_destroy3d(sfd_handle3d h) {
    if (!h) return SFD_EINVAL;
    delete h;
    return SFD_OK;
}

const char *sfd_last_error3d(sfd_handle3d h) { return h ? h->err.c_str() : g_create_err3d.c_str(); }

int sfd_set_model3d(sfd_handle3d h, const float *vel, const float *vel_bwd) {
    if (!h || !vel) return SFD_EINVAL;
    return abi_call(&h->err, [&]() {
        const size_t n = (size_t)h->g.nz * h->g.nx * h->g.ny;
        for (int m = 0; m < 2; m++) {
            const float *v = m == 0 ? vel : vel_bwd;
            if (!v) continue;
            for (size_t i = 0; i < n; i++)
                if (!(v[i] > 0.f) || !std::isfinite(v[i])) throw EngineError(SFD_EINVAL, "velocity must be finite and > 0");
            vel_check3(h->g, v, n);
        }
        h->vel_h.assign(vel, vel + n);
        if (!h->d_vel) HIPCHK(hipMalloc(&h->d_vel, n * sizeof(float)));
        HIPCHK(hipMemcpy(h->d_vel, vel, n * sizeof(float), hipMemcpyHostToDevice));
        if (vel_bwd) {
            h->velb_h.assign(vel_bwd, vel_bwd + n);
            if (!h->d_velb) HIPCHK(hipMalloc(&h->d_velb, n * sizeof(float)));
            HIPCHK(hipMemcpy(h->d_velb, vel_bwd, n * sizeof(float), hipMemcpyHostToDevice));
        } else {
            h->velb_h.clear();
            if (h->d_velb) (void)hipFree(h->d_velb);
            h->d_velb = nullptr;
        }
    });
}

int sfd_set_density3d(sfd_handle3d h, const float *rho) {
    if (!h) return SFD_EINVAL;
    return abi_call(&h->err, [&]() {
        if (!rho) {
            h->rho_h.clear();
            return;
        }
        const size_t n = (size_t)h->g.nz * h->g.nx * h->g.ny;
        for (size_t i = 0; i < n; i++)
            if (!(rho[i] > 0.f) || !std::isfinite(rho[i])) throw EngineError(SFD_EINVAL, "density must be finite and > 0");
        if (!h->d_rho) HIPCHK(hipMalloc(&h->d_rho, n * sizeof(float)));
        HIPCHK(hipMemcpy(h->d_rho, rho, n * sizeof(float), hipMemcpyHostToDevice));
        h->rho_h.assign(rho, rho + n);
    });
}

int sfd_forward3d(sfd_handle3d h, const sfd_shot3d *shots, int nshots, float *traces_out, int extract_kind) {
    if (!h || !shots || nshots < 1 || !traces_out) return SFD_EINVAL;
    if (extract_kind < 0 || extract_kind > 3) return SFD_EINVAL;
    return abi_call(&h->err, [&]() {
        if (h->vel_h.empty()) throw EngineError(SFD_ESTATE, "no model: call sfd_set_model3d first");
        size_t off = 0;
        for (int s = 0; s < nshots; s++) {
            forward3(*h, shots[s], traces_out + off, extract_kind);
            off += (size_t)shots[s].nt * shots[s].ntr;
        }
    });
}

int sfd_rtm3d(sfd_handle3d h, const sfd_shot3d *shots, int nshots, const float *traces_in, float *image,
              float *illum) {
    if (!h || !shots || nshots < 1 || !traces_in || (!image && !illum)) return SFD_EINVAL;
    return abi_call(&h->err, [&]() {
        if (h->vel_h.empty()) throw EngineError(SFD_ESTATE, "no model: call sfd_set_model3d first");
        size_t off = 0;
        for (int s = 0; s < nshots; s++) {
            rtm3(*h, shots[s], traces_in + off, image, illum);
            off += (size_t)shots[s].nt * shots[s].ntr;
        }
    });
}

int sfd_lsrtm3d(sfd_handle3d h, const sfd_shot3d *shots, int nshots, const float *traces_in, float *image,
                float *illum) {
    if (!h || !shots || nshots < 1 || !traces_in || (!image && !illum)) return SFD_EINVAL;
    return abi_call(&h->err, [&]() {
        if (h->vel_h.empty()) throw EngineError(SFD_ESTATE, "no model: call sfd_set_model3d first");
        if (!h->rho_h.empty()) throw EngineError(SFD_EUNSUPPORTED, "LSRTM runs with constant density (LSRTM_to_memory)");
        if (h->g.snapi < 1) throw EngineError(SFD_EINVAL, "snapi must be >= 1");
        size_t off = 0;
        for (int s = 0; s < nshots; s++) {
            lsrtm3(*h, shots[s], traces_in + off, image, illum);
            off += (size_t)shots[s].nt * shots[s].ntr;
        }
    });
}

int sfd_born3d(sfd_handle3d h, const sfd_shot3d *shots, int nshots, const float *refl, float *traces_out) {
    if (!h || !shots || nshots < 1 || !refl || !traces_out) return SFD_EINVAL;
    return abi_call(&h->err, [&]() {
        if (h->vel_h.empty()) throw EngineError(SFD_ESTATE, "no model: call sfd_set_model3d first");
        if (!h->rho_h.empty()) throw EngineError(SFD_EUNSUPPORTED, "Born runs with constant density (BORN_to_memory)");
        if (h->g.snapi < 1) throw EngineError(SFD_EINVAL, "snapi must be >= 1");
        const size_t n = (size_t)h->g.nz * h->g.nx * h->g.ny;
        float *d_refl = nullptr;
        HIPCHK(hipMalloc(&d_refl, n * 4));
        std::unique_ptr<float, decltype(&hipFree)> guard(d_refl, &hipFree);
        HIPCHK(hipMemcpy(d_refl, refl, n * 4, hipMemcpyHostToDevice));
        size_t off = 0;
        for (int s = 0; s < nshots; s++) {
            born3(*h, shots[s], d_refl, traces_out + off);
            off += (size_t)shots[s].nt * shots[s].ntr;
        }
    });
}

int sfd_afwi_gradient3d(sfd_handle3d h, const sfd_shot3d *shots, int nshots, const float *dobs, const float *mute,
                        const sfd_afwi_opts *opt, double *misfit, float *grad, float *illum) {
    if (!h || !shots || nshots < 1 || !dobs || !opt || !misfit || !grad) return SFD_EINVAL;
    return abi_call(&h->err, [&]() {
        const sfd_grid3d &g = h->g;
        if (h->vel_h.empty()) throw EngineError(SFD_ESTATE, "no model: call sfd_set_model3d first");
        if (g.snapi < 1) throw EngineError(SFD_EINVAL, "snapi must be >= 1");
        if (opt->snap_mode != SFD_SNAP_FP32) throw EngineError(SFD_EUNSUPPORTED, "the 3D path keeps FP32 snapshots");
        const int norm = opt->norm == 0 ? SFD_NORM_L2 : opt->norm;
        if (norm != SFD_NORM_L2 && norm != SFD_NORM_L1 && norm != SFD_NORM_HUBER && norm != SFD_NORM_CAUCHY &&
            norm != SFD_NORM_HYPERBOLIC)
            throw EngineError(SFD_EINVAL, "unknown norm (SFD_NORM_*)");
        if (!std::isfinite(opt->norm_thresh)) throw EngineError(SFD_EINVAL, "norm_thresh is not finite");
        if (opt->vprho_param != 0 && opt->vprho_param != 1) throw EngineError(SFD_EINVAL, "vprho_param must be 0 or 1");
        const size_t N = (size_t)g.nz * g.nx * g.ny;
        const bool rho = !h->rho_h.empty();
        std::fill(grad, grad + N * (rho ? 2 : 1), 0.f);
        if (illum) std::fill(illum, illum + N, 0.f);
        double f = 0;
        size_t off = 0;
        for (int s = 0; s < nshots; s++) {
            afwi3(*h, shots[s], dobs + off, mute ? mute + off : nullptr, norm, opt->norm_thresh, grad, illum, f);
            off += (size_t)shots[s].nt * shots[s].ntr;
        }
        if (!std::isfinite(f)) throw EngineError(SFD_ENOTFINITE, "misfit is not finite");
        if (rho && opt->vprho_param == 1)   // ExtractPadModel.f90:643-646 on the model's vel and rho
            for (size_t q = 0; q < N; q++) {
                const float v = h->vel_h[q], r = h->rho_h[q];
                grad[q] = grad[q] - (r / v) * grad[N + q];
                grad[N + q] = (1.f / v) * grad[N + q];
            }
        *misfit = f;
    });
}

}  // extern "C"
