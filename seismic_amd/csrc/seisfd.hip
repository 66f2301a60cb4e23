// seisfd.hip — host side of the MI355X acoustic FD engine and its C-ABI
// (include/seisfd.h).
//
// The reference's per-shot Fortran drivers (progs/modeling, progs/rtm,
// progs/afwi) become batched calls: every shot of a call is windowed, padded
// and planned on the host once, then all shots advance together, one fused
// step kernel launch per time step (fd_kernels.h), captured into HIP graphs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/seisfd.h"
#include "fd_kernels.h"
#include "fwi_kernels.h"
#include "bandpass_kernels.h"
#include "engine.h"
#include "fd_step2.h"
#include "adjoint_kernels.h"
#include "common.h"
#include "cnn_api.h"
#include "lbfgs_api.h"

using namespace sfd;

// ---------------------------------------------------------------------------
// Reference arithmetic needed on the host (FP32, same operation order).

static inline int nint_f(float x) { return (int)roundf(x); }   // Fortran nint

// ModelSpaceTypes.f90:330-337 find_i_d
static inline void find_i_d(int *ic, float *dc, float c, float o, float d) {
    *ic = nint_f((c - o) / d) + 1;
    *dc = c - (o + (float)(*ic - 1) * d);
}

// Interpolate.f90:15-42 mksinc
void sfd_host_mksinc(float *sinc, int lsinc, float shift) {
    const float pi = 3.141592654f;
    float fsum = 0.f, fact, pos;
    for (int n = 1; n <= lsinc; n++) {
        if (n <= (int)((float)(lsinc / 2 + 1) + shift))
            fact = ((float)(n - 1) - shift) / (float)(lsinc / 2);
        else
            fact = ((float)(lsinc - n) + shift) / (float)(lsinc / 2);
        if (fact <= 0.f) fact = 0.f;
        pos = pi * ((float)(n - (lsinc / 2 + 1)) - shift);
        if (fabsf(pos) > 0.005f)
            sinc[n - 1] = fact * sinf(pos) / pos;
        else
            sinc[n - 1] = fact * (1.f - pos * pos / 6.f);
        fsum = fsum + sinc[n - 1];
    }
    for (int n = 1; n <= lsinc; n++) sinc[n - 1] = sinc[n - 1] / fsum;
}

// Taper.f90:14-44 compute_taper / cos_sin_taper (also the 3D engine's, per axis)
void sfd_host_taper(float *t, int n) {
    const float pi = 3.14159265359f;
    const int nt = std::min(n / 2, 10);
    const int m = std::min(nt, n);
    for (int i = 0; i < n; i++) t[i] = 1.f;
    for (int i = 1; i <= m; i++) {
        const float s = sinf(pi * (float)(i - 1) / (float)(2 * (m - 1)));
        t[i - 1] = s * s;
    }
    const int b = std::max(1, n - nt);
    for (int i = b; i <= n; i++) {
        const float c = cosf(pi * (float)(i - b) / (float)(2 * nt));
        t[i - 1] = c * c;
    }
}

// ExtractPadModel.f90:19-85 extract_coord_source_receiver_patch + nxw (:183)
static void shot_window(const sfd_grid &g, const sfd_shot &s, float *oxw, int *nxw) {
    float sx = s.sx, gx1 = s.rx[0], gxn = s.rx[0];
    for (int i = 1; i < s.ntr; i++) {
        gx1 = std::min(gx1, s.rx[i]);
        gxn = std::max(gxn, s.rx[i]);
    }
    const float minx = std::min(sx, gx1), maxx = std::max(sx, gxn);
    const float mx1 = (gx1 + sx) / 2, mxn = (gxn + sx) / 2;
    const float cx = mx1 + fabsf(mxn - mx1) / 2;
    *oxw = std::min(minx, cx - g.aperture_x / 2);
    const float endx = std::max(maxx, cx + g.aperture_x / 2);
    *nxw = nint_f((endx - *oxw) / g.dx) + 1;
}

// ---------------------------------------------------------------------------

void DevBuf::reserve(size_t n) {
    if (n <= cap) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    if (hipMalloc(&p, n) != hipSuccess) {
        p = nullptr;
        throw EngineError(SFD_ENOMEM, "device allocation of " + std::to_string(n) + " bytes failed");
    }
    cap = n;
}
DevBuf::~DevBuf() {
    if (p) (void)hipFree(p);
}
void HostBuf::reserve(size_t n) {
    if (n <= cap) return;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    if (hipHostMalloc(&p, n) != hipSuccess) {
        p = nullptr;
        throw EngineError(SFD_ENOMEM, "pinned host allocation of " + std::to_string(n) + " bytes failed");
    }
    cap = n;
}
HostBuf::~HostBuf() {
    if (p) (void)hipHostFree(p);
}

static void check_grid(const sfd_grid &g) {
    if (g.nz < 2 || g.nx < 1 || !(g.dz > 0) || !(g.dx > 0) || !(g.dt > 0))
        throw EngineError(SFD_EINVAL, "bad grid dimensions or spacing");
    if (g.ntaper < 10) throw EngineError(SFD_EINVAL, "ntaper must be >= 10 (Higdon bands and injection must not overlap)");
    if (g.lsinc != 7) throw EngineError(SFD_EUNSUPPORTED, "only lsinc=7 is implemented");
    if (g.snapi < 1) throw EngineError(SFD_EINVAL, "snapi must be >= 1");
    if (!(g.fmax >= 0.f) || !std::isfinite(g.fmax)) throw EngineError(SFD_EINVAL, "fmax must be >= 0 (0: 30 Hz)");
}

// Rows of the FD region above model row 1 (1 - nmin1): read_window_vel's
// bounds (ExtractPadModel.f90:228-233) — a ghost source needs room for its
// mirror image above the surface, a free surface keeps lsinc rows (nmin1 =
// 1 - lsinc).
static int top_rows(const sfd_grid &g) {
    if (g.surf_type > 0) return g.lsinc;
    if (g.shot_type > 0) return g.ntaper + g.lsinc / 2 + 2;
    return g.ntaper;
}

// ReadSouVelRho.f90:603-752 vel_check, 2D: the Courant stop (:629-675,
// vmax dt / min(dz,dx) > 0.45) and the sampling stop (:688-752: fewer than
// 2 cells of max(dz,dx) per shortest wavelength vmin / fmax).
static void vel_check(const sfd_grid &g, const float *v, size_t n) {
    float vmax = v[0], vmin = v[0];
    for (size_t i = 1; i < n; i++) {
        vmax = std::max(vmax, v[i]);
        vmin = std::min(vmin, v[i]);
    }
    const float stab = vmax * g.dt / std::min(g.dz, g.dx);
    char b[200];
    if (stab > 0.45f) {
        snprintf(b, sizeof b, "Courant number %.6g > 0.45: unstable (largest stable dt %.6g)", stab,
                 0.45f / stab * g.dt);
        throw EngineError(SFD_EUNSTABLE, b);
    }
    const float fmax = g.fmax > 0.f ? g.fmax : 30.f;
    const float samplerate = vmin / fmax / std::max(g.dz, g.dx);
    if (samplerate < 2.0f) {
        snprintf(b, sizeof b, "wavefield not adequately sampled: %.6g cells per shortest wavelength (vmin %.6g, "
                 "fmax %.6g Hz); lower fmax to %.6g or refine the grid", samplerate, vmin, fmax,
                 samplerate / 3.0f * fmax);
        throw EngineError(SFD_EALIASED, b);
    }
}

Engine::Engine(const sfd_grid &grid) : g(grid) {
    check_grid(g);
    HIPCHK(hipStreamCreate(&stream));
    HIPCHK(hipEventCreate(&ev0));
    HIPCHK(hipEventCreate(&ev1));
    HIPCHK(hipEventCreate(&ev2));
}

Engine::~Engine() {
    for (auto &e : kev) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (ev2) (void)hipEventDestroy(ev2);
    if (d_vel) (void)hipFree(d_vel);
    if (d_velb) (void)hipFree(d_velb);
    if (d_rho) (void)hipFree(d_rho);
}

void Engine::set_model(const float *vel, const float *vel_bwd) {
    const size_t n = (size_t)g.nz * g.nx;
    if (!vel) throw EngineError(SFD_EINVAL, "vel is NULL");
    for (size_t i = 0; i < n; i++)
        if (!(vel[i] > 0.f) || !std::isfinite(vel[i])) throw EngineError(SFD_EINVAL, "velocity must be finite and > 0");
    vel_check(g, vel, n);
    vel_h.assign(vel, vel + n);
    model_version++;
    if (!d_vel) HIPCHK(hipMalloc(&d_vel, n * sizeof(float)));
    HIPCHK(hipMemcpy(d_vel, vel, n * sizeof(float), hipMemcpyHostToDevice));
    if (vel_bwd) {
        vel_check(g, vel_bwd, n);
        velb_h.assign(vel_bwd, vel_bwd + n);
        if (!d_velb) HIPCHK(hipMalloc(&d_velb, n * sizeof(float)));
        HIPCHK(hipMemcpy(d_velb, vel_bwd, n * sizeof(float), hipMemcpyHostToDevice));
    } else {
        velb_h.clear();
        if (d_velb) (void)hipFree(d_velb);
        d_velb = nullptr;
    }
    have_model = true;
}

// Free-surface elevation per model column (the reference's elev%elev, which
// its drivers set to 0: progs/rtm/Acoustic_RTM_sep.f90:118-119); NULL: 0.
void Engine::set_elevation(const float *elev) {
    if (!elev) {
        elev_h.clear();
    } else {
        for (int i = 0; i < g.nx; i++)
            if (!std::isfinite(elev[i])) throw EngineError(SFD_EINVAL, "elevation must be finite");
        elev_h.assign(elev, elev + g.nx);
    }
    model_version++;   // plans hold the surface rows
}

// Variable density (genpar%withRho with the model's rho, ReadSouVelRho.f90:
// 504-516): rho per model cell, NULL for constant density.
void Engine::set_density(const float *rho) {
    const size_t n = (size_t)g.nz * g.nx;
    if (!rho) {
        rho_h.clear();
        if (d_rho) (void)hipFree(d_rho);
        d_rho = nullptr;
    } else {
        for (size_t i = 0; i < n; i++)
            if (!(rho[i] > 0.f) || !std::isfinite(rho[i])) throw EngineError(SFD_EINVAL, "density must be finite and > 0");
        rho_h.assign(rho, rho + n);
        if (!d_rho) HIPCHK(hipMalloc(&d_rho, n * sizeof(float)));
        HIPCHK(hipMemcpy(d_rho, rho, n * sizeof(float), hipMemcpyHostToDevice));
    }
    model_version++;
}

// The source as the propagation injects it: with variable density readsou
// shifts it half a time step forward with the lsinc-point sinc, renormalised
// where the interpolator runs off the trace (ReadSouVelRho.f90:328-349).
const float *Engine::source_wavelet(const sfd_shot &sh, int k) {
    if (!has_rho()) return sh.wavelet;
    if ((int)src_shifted.size() <= k) src_shifted.resize(k + 1);
    std::vector<float> &o = src_shifted[k];
    o.assign(sh.nt, 0.f);
    float w[64];
    const int h = g.lsinc / 2;
    sfd_host_mksinc(w, g.lsinc, 0.5f);
    for (int it = 1; it <= sh.nt; it++) {
        float fscale = 1.f, fdum = 0.f;
        for (int l = -h; l <= h; l++) {
            if (it + l >= 1 && it + l <= sh.nt) fdum = fdum + w[h + l] * sh.wavelet[it + l - 1];
            else fscale = fscale - w[h + l];
        }
        o[it - 1] = fdum / fscale;
    }
    return o.data();
}

// Padded density at a Fortran window cell, as padded_vel.
float Engine::padded_rho(const ShotPlan &p, int icz, int icx) const {
    int wz = std::min(std::max(icz, 1), g.nz) - 1;
    int wx = std::min(std::max(icx, 1), p.nxw) - 1;
    return rho_h[(size_t)p.colsrc[wx] * g.nz + wz];
}

// Padded velocity at a Fortran window cell (icz, icx) (1-based, may lie in
// the pad): the edge-replicated value model_pad would hold there.
float Engine::padded_vel(const std::vector<float> &vh, const ShotPlan &p, int icz, int icx) const {
    int wz = std::min(std::max(icz, 1), g.nz) - 1;
    int wx = std::min(std::max(icx, 1), p.nxw) - 1;
    return vh[(size_t)p.colsrc[wx] * g.nz + wz];
}

// Runs f(s) for s in [0, n) on up to 16 host threads: the shots' plans are
// independent (every planner is const and writes only its own plan).  The
// first failure in shot order is rethrown after all threads have joined.
template <class F>
static void for_shots(int n, F f) {
    const int hw = (int)std::thread::hardware_concurrency();
    const int nt = std::max(1, std::min(std::min(n, 16), hw > 0 ? hw : 1));
    if (nt <= 1) {
        for (int s = 0; s < n; s++) f(s);
        return;
    }
    std::vector<std::exception_ptr> err(n);
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; t++)
        pool.emplace_back([&, t]() {
            for (int s = t; s < n; s += nt) {
                try {
                    f(s);
                } catch (...) {
                    err[s] = std::current_exception();
                }
            }
        });
    for (auto &th : pool) th.join();
    for (auto &e : err)
        if (e) std::rethrow_exception(e);
}

void Engine::plan_shot(const sfd_shot &s, ShotPlan &p) const {
    if (s.nt < 1 || s.ntr < 1 || !s.rz || !s.rx) throw EngineError(SFD_EINVAL, "shot needs nt >= 1 and >= 1 receiver");
    shot_window(g, s, &p.oxw, &p.nxw);
    if (p.nxw < 1) throw EngineError(SFD_EINVAL, "empty shot window");
    p.nt = s.nt;
    p.ntr = s.ntr;
    p.ntsnap = s.nt / g.snapi;
    p.ntop = top_rows(g);
    p.NZ = g.nz + p.ntop + g.ntaper + 8;
    p.NX = p.nxw + 2 * g.ntaper + 8;
    p.i0 = 4 + g.ntaper;
    p.iz0 = 4 + p.ntop;
    const int rz_ = (p.NZ - 1) % tz;
    p.zsh = rz_ >= 9 ? 0 : 9 - rz_;
    {
        // Where it adds no tile row, shift the rows so that the shallowest
        // injection footprint (source or receivers, 3 rows above the centre)
        // lies below the top tile row: the top Higdon band and the injection
        // staging then run in different tiles, and a launch ends with its
        // slowest tile.  The top tile keeps rows 0..9 (band + the two rows
        // its recursion reads), the bottom tile the last 10 rows.
        float zmin = s.sz;
        for (int r = 0; r < s.ntr; r++) zmin = std::min(zmin, s.rz[r]);
        const int rmin = p.iz0 + (int)std::floor((zmin - g.oz) / g.dz) - 4;
        const int want = tz - rmin;   // the smallest shift that puts row rmin in tile row 1
        if (want > p.zsh && want <= tz - 10 && (p.NZ - 1 + want) % tz >= 9 &&
            (p.NZ + want + tz - 1) / tz == (p.NZ + p.zsh + tz - 1) / tz)
            p.zsh = want;
    }
    const int rx_ = (p.NX - 1) % tx;
    p.xsh = rx_ >= 9 ? 0 : 9 - rx_;
    p.ntz = (p.NZ + p.zsh + tz - 1) / tz;
    p.ntx = (p.NX + p.xsh + tx - 1) / tx;
    p.zoff = p.zsh;
    p.LDZ = ((p.zoff + p.NZ + 31) / 32) * 32;
    // nearest-neighbour window columns (ExtractPadModel.f90:516-522)
    p.colsrc.resize(p.nxw);
    for (int j = 1; j <= p.nxw; j++) {
        const float cxx = (float)(j - 1) * g.dx + p.oxw;
        int ix = nint_f((cxx - g.ox) / g.dx) + 1;
        p.colmap_raw.push_back(ix);
        ix = std::min(std::max(ix, 1), g.nx);
        p.colsrc[j - 1] = ix - 1;
    }
    p.taper.resize(p.nxw);
    sfd_host_taper(p.taper.data(), p.nxw);
    // Free surface (ModelSpace_elevation_parameters, ModelSpaceTypes.f90:128-141):
    // per FD column, ielev_z / delev_z of its elevation by find_i_d against
    // the depth axis, and the sinc of delev_z / dz that an off-grid surface
    // interpolates with (Boundary.f90:1636).  FD columns outside the window
    // take the elevation of the nearest window column.
    p.fs_row.clear();
    p.fs_w.clear();
    p.snap_mask = 0;
    if (g.surf_type != 0) {
        p.fs_row.assign(p.NX, 0);
        p.fs_w.assign((size_t)8 * p.NX, 0.f);
        for (int ix = 4; ix <= p.NX - 5; ix++) {
            const int wx = std::min(std::max(ix - p.i0, 0), p.nxw - 1);
            const float e = elev_h.empty() ? 0.f : elev_h[p.colsrc[wx]];
            int ie;
            float de;
            find_i_d(&ie, &de, e, g.oz, g.dz);
            if (ie < 1 || ie > g.nz)
                throw EngineError(SFD_EINVAL, "free-surface elevation must lie within the model depth range");
            p.fs_row[ix] = ie - 1 + p.iz0;
            if (de != 0.f) {
                sfd_host_mksinc(&p.fs_w[(size_t)8 * ix], 7, de / g.dz);
                p.fs_w[(size_t)8 * ix + 7] = 1.f;
            }
            if (ie > 1 && ix >= p.i0 && ix < p.i0 + p.nxw) p.snap_mask = 1;
        }
    }
}

// Tile of a padded cell, and its local cell index.
static inline void cell_tile(const ShotPlan &p, int TZ, int TX, int iz, int ix, int *t, int *local) {
    const int tzi = (iz + p.zsh) / TZ, txi = (ix + p.xsh) / TX;
    const int z0 = tzi * TZ - p.zsh, x0 = txi * TX - p.xsh;
    *t = tzi + p.ntz * txi;
    *local = (iz - z0) + TZ * (ix - x0);
}

// Region plan of the two-step kernel from a per-tile plan: for every tile,
// the cells of its own and its 8 neighbours' plans that fall inside its
// 72x40 region (tile + 4-cell ring), with their contributions in the same
// order, re-indexed region-locally and against a per-region trace table.
static RegionPlan region_plan(const ShotPlan &p, const std::vector<int> &tile, const std::vector<int> &cell,
                              const std::vector<int> &cptr, const std::vector<float> &coef,
                              const std::vector<int> &tptr, const std::vector<int> &ttr) {
    const int TZ = B2Z, TX = B2X, ntile = p.ntz * p.ntx;
    RegionPlan r;
    r.tile.assign(ntile + 1, 0);
    r.tptr.assign(ntile + 1, 0);
    std::vector<int> local;
    std::vector<int> used;
    for (int t = 0; t < ntile; t++) {
        const int tzi = t % p.ntz, txi = t / p.ntz;
        const int z0 = tzi * TZ - p.zsh, x0 = txi * TX - p.xsh;
        for (int u : used) local[u] = -1;
        used.clear();
        for (int dx = -1; dx <= 1; dx++)
            for (int dz = -1; dz <= 1; dz++) {
                const int nzi = tzi + dz, nxi = txi + dx;
                if (nzi < 0 || nzi >= p.ntz || nxi < 0 || nxi >= p.ntx) continue;
                const int n = nzi + p.ntz * nxi;
                const int nz0 = nzi * TZ - p.zsh, nx0 = nxi * TX - p.xsh;
                for (int k = tile[n]; k < tile[n + 1]; k++) {
                    const int lz = cell[k] % TZ, lx = cell[k] / TZ;
                    const int ez = nz0 + lz - (z0 - 4), ex = nx0 + lx - (x0 - 4);
                    if (ez < 0 || ez >= R1Z || ex < 0 || ex >= R1X) continue;
                    r.cell.push_back(ez + R1Z * ex);
                    r.cptr.push_back((int)(r.coef.size() / 4));
                    r.tile[t + 1]++;
                    for (int q = cptr[k]; q < cptr[k + 1]; q++) {
                        int li;
                        memcpy(&li, &coef[4 * q + 3], 4);
                        const int gtr = ttr[tptr[n] + li];
                        if (gtr >= (int)local.size()) local.resize(gtr + 1, -1);
                        if (local[gtr] < 0) {
                            local[gtr] = (int)used.size();
                            used.push_back(gtr);
                            r.ttr.push_back(gtr);
                            r.tptr[t + 1]++;
                        }
                        int nl = local[gtr];
                        float nlf;
                        memcpy(&nlf, &nl, 4);
                        r.coef.insert(r.coef.end(), {coef[4 * q], coef[4 * q + 1], coef[4 * q + 2], nlf});
                    }
                }
            }
        if ((int)used.size() > kMaxTileTraces2) r.ok = false;
    }
    r.cptr.push_back((int)(r.coef.size() / 4));
    for (int t = 0; t < ntile; t++) {
        r.tile[t + 1] += r.tile[t];
        r.tptr[t + 1] += r.tptr[t];
    }
    if (r.coef.size() / 4 >= (1u << 23)) r.ok = false;
    return r;
}

// [lo, hi): the tiles t with tab[t+1] > tab[t] (empty: 0, 0)
static void tile_range(const std::vector<int> &tab, int &lo, int &hi) {
    lo = hi = 0;
    for (int t = 0; t + 1 < (int)tab.size(); t++)
        if (tab[t + 1] > tab[t]) {
            if (hi == 0) lo = t;
            hi = t + 1;
        }
}

// The step kernel's form of an injection plan: per tile a descriptor of 8
// ints {first contribution, end, first trace, end, cell-map offset, 0, 0, 0}
// and, for every injecting tile, a dense map over its cells of
// (start - first) << 8 | count.
static void inj_dense(int ntile, int ncell, const std::vector<int> &tile, const std::vector<int> &cell,
                      const std::vector<int> &cptr, const std::vector<int> &tptr, const std::vector<int> &ttr,
                      std::vector<int> &desc, std::vector<int> &map) {
    desc.assign((size_t)ntile * 8, 0);
    map.clear();
    for (int t = 0; t < ntile && t + 1 < (int)tile.size(); t++) {
        const int c0 = tile[t], c1 = tile[t + 1];
        if (c1 <= c0) continue;
        int *dt = &desc[(size_t)t * 8];
        dt[0] = cptr[c0];
        dt[1] = cptr[c1];
        dt[2] = tptr[t];
        dt[3] = tptr[t + 1];
        dt[4] = (int)map.size();
        dt[5] = ttr[dt[2]];   // consecutive global traces: the first; else -1
        for (int j = dt[2] + 1; j < dt[3]; j++)
            if (ttr[j] != ttr[j - 1] + 1) dt[5] = -1;
        map.resize(map.size() + ncell, 0);
        int *m = &map[(size_t)dt[4]];
        for (int c = c0; c < c1; c++) m[cell[c]] = ((cptr[c] - dt[0]) << 8) | (cptr[c + 1] - cptr[c]);
    }
}

void InjSet::apply(sfd::ShotDev &S) const {
    S.inj_tile = tile; S.inj_cell = cell; S.inj_cptr = cptr; S.inj_coef = coef; S.inj_tptr = tptr; S.inj_ttr = ttr;
    S.inj_lo = lo; S.inj_hi = hi; S.inj_desc = desc; S.inj_map = map;
    S.inj2_tile = tile2; S.inj2_cell = cell2; S.inj2_cptr = cptr2; S.inj2_coef = coef2; S.inj2_tptr = tptr2;
    S.inj2_ttr = ttr2;
}

// Injection plan: kind 1 = Injection_sinc (gain v^2, Injection.f90:151-194),
// 2 = Injection_simple_noomp (v^2, :63-104), 3 = Injection_LSRTM_sinc (1/(dz dx), :239-272).
void Engine::plan_injection(ShotPlan &p, const std::vector<float> &vh, int nsrc, const float *sz_, const float *sx_,
                            int kind) const {
    struct C { int iz, ix, tr; float a, wz, wx; };
    std::vector<C> cs;
    // Ghosts: the sinc injections test shot_type whichever traces they inject
    // (Injection.f90:432); the simple one type_inject, which is rec_type in
    // the reverse-time passes that use it (:74-78, :583).  A ghost subtracts
    // the same term at the mirror cell M - (iz - i) with the weight of -i,
    // right after the direct term: a negated gain gives that product exactly.
    if ((kind == 4 || kind == 5) && ghosts_or_surface())
        throw EngineError(SFD_EUNSUPPORTED, "the exact adjoints have no free surface or ghosts");
    const bool ghost = (kind == 2) ? g.rec_type != 0 : g.shot_type != 0;
    const int M = 2 * p.iz0 - 2;
    cs.reserve((size_t)nsrc * (ghost ? 98 : 49));
    float wzs[7], wxs[7];
    for (int k = 0; k < nsrc; k++) {
        int icz, icx;
        float dcz, dcx;
        find_i_d(&icz, &dcz, sz_[k], g.oz, g.dz);
        find_i_d(&icx, &dcx, sx_[k], p.oxw, g.dx);
        const int iz = icz - 1 + p.iz0, ix = icx - 1 + p.i0;
        const int hz = (kind == 2 || kind == 5) ? 0 : 3;
        if (iz - hz < 4 || iz + hz > p.NZ - 5 || ix - hz < 4 || ix + hz > p.NX - 5)
            throw EngineError(SFD_EINVAL, "source/receiver footprint leaves the FD region");
        if (ghost && (M - iz - hz < 4 || M - iz + hz > p.NZ - 5))
            throw EngineError(SFD_EINVAL, "ghost source footprint leaves the FD region");
        float a;
        if (kind == 3) {
            a = 1.f / (g.dz * g.dx);
        } else if (kind == 4 || kind == 5) {
            a = 1.f;
        } else {
            const float v = padded_vel(vh, p, icz, icx);
            a = v * v;
            if (has_rho()) a = a * padded_rho(p, icz, icx);   // Injection.f90:172-176
        }
        if (kind == 2 || kind == 5) {
            cs.push_back({iz, ix, k, a, 1.f, 1.f});
            if (ghost) cs.push_back({M - iz, ix, k, -a, 1.f, 1.f});
            continue;
        }
        sfd_host_mksinc(wzs, 7, dcz / g.dz);
        sfd_host_mksinc(wxs, 7, dcx / g.dx);
        for (int j = -3; j <= 3; j++)
            for (int i = -3; i <= 3; i++) {
                cs.push_back({iz + i, ix + j, k, a, wzs[3 + i], wxs[3 + j]});
                if (ghost) cs.push_back({M - (iz - i), ix + j, k, -a, wzs[3 - i], wxs[3 + j]});
            }
    }
    // group by cell, keeping trace order within a cell (stable sort)
    const int ntile = p.ntz * p.ntx;
    std::vector<std::pair<long, int>> key(cs.size());
    for (size_t i = 0; i < cs.size(); i++) {
        int t, loc;
        cell_tile(p, tz, tx, cs[i].iz, cs[i].ix, &t, &loc);
        key[i] = {(long)t * (long)(tz * tx) + loc, (int)i};
    }
    std::stable_sort(key.begin(), key.end(), [](const std::pair<long, int> &a, const std::pair<long, int> &b) {
        return a.first < b.first;
    });
    p.inj_tile.assign(ntile + 1, 0);
    p.inj_cell.clear(); p.inj_cptr.clear(); p.inj_coef.clear();
    p.inj_tptr.assign(ntile + 1, 0);
    p.inj_ttr.clear();
    long prev = -1;
    int cur_tile = -1;
    std::vector<int> local;   // trace -> tile-local index (-1: not yet used in this tile)
    std::vector<int> used;
    local.assign(std::max(nsrc, 1), -1);
    for (size_t i = 0; i < key.size(); i++) {
        const C &c = cs[key[i].second];
        const int t = (int)(key[i].first / (tz * tx));
        if (t != cur_tile) {
            for (int u : used) local[u] = -1;
            used.clear();
            cur_tile = t;
        }
        if (key[i].first != prev) {
            p.inj_cell.push_back((int)(key[i].first % (tz * tx)));
            p.inj_cptr.push_back((int)(p.inj_coef.size() / 4));
            p.inj_tile[t + 1]++;
            prev = key[i].first;
        }
        if (local[c.tr] < 0) {
            local[c.tr] = (int)used.size();
            used.push_back(c.tr);
            p.inj_ttr.push_back(c.tr);
            p.inj_tptr[t + 1]++;
            if ((int)used.size() > kMaxTileTraces)
                throw EngineError(SFD_EUNSUPPORTED, "more than 1024 distinct injected traces in one tile");
        }
        int li = local[c.tr];
        float lif;
        memcpy(&lif, &li, 4);
        p.inj_coef.insert(p.inj_coef.end(), {c.a, c.wz, c.wx, lif});
    }
    p.inj_cptr.push_back((int)(p.inj_coef.size() / 4));
    for (int t = 0; t < ntile; t++) {
        p.inj_tile[t + 1] += p.inj_tile[t];
        p.inj_tptr[t + 1] += p.inj_tptr[t];
    }
    for (size_t c = 0; c + 1 < p.inj_cptr.size(); c++)
        if (p.inj_cptr[c + 1] - p.inj_cptr[c] > 255) throw EngineError(SFD_EUNSUPPORTED, "more than 255 injections in one cell");
    if (p.inj_coef.size() / 4 >= (1u << 23)) throw EngineError(SFD_EUNSUPPORTED, "injection plan too large");
}

// Extraction plan (Extraction.f90:71-246).
void Engine::plan_extraction(ShotPlan &p, const std::vector<float> &vh, const sfd_shot &s, int kind) const {
    const int ntile = p.ntz * p.ntx;
    struct R { int t, loc, r; float w[16]; };
    std::vector<R> rs;
    for (int r = 0; r < s.ntr; r++) {
        int icz, icx;
        float dcz, dcx;
        find_i_d(&icz, &dcz, s.rz[r], g.oz, g.dz);
        find_i_d(&icx, &dcx, s.rx[r], p.oxw, g.dx);
        const int iz = icz - 1 + p.iz0, ix = icx - 1 + p.i0;
        if (iz - 3 < 0 || iz + 3 >= p.NZ || ix - 3 < 0 || ix + 3 >= p.NX)
            throw EngineError(SFD_EINVAL, "receiver outside the padded grid");
        if (g.rec_type != 0 && (2 * p.iz0 - 2 - iz - 3 < 0 || 2 * p.iz0 - 2 - iz + 3 >= p.NZ))
            throw EngineError(SFD_EINVAL, "receiver ghost outside the padded grid");
        R e;
        memset(&e, 0, sizeof e);
        cell_tile(p, tz, tx, iz, ix, &e.t, &e.loc);
        e.r = r;
        if (kind == EXT_SINC_V2) {
            const float v = padded_vel(vh, p, icz, icx);
            e.w[0] = v * v;
        } else if (kind == EXT_SINC_LSRTM) {
            e.w[0] = 1.f / (g.dz * g.dx);
        } else {
            e.w[0] = 1.f;
        }
        if (kind != EXT_NEAREST) {
            sfd_host_mksinc(e.w + 1, 7, dcz / g.dz);
            sfd_host_mksinc(e.w + 8, 7, dcx / g.dx);
        }
        rs.push_back(e);
    }
    std::stable_sort(rs.begin(), rs.end(), [](const R &a, const R &b) { return a.t < b.t; });
    p.ext_tile.assign(ntile + 1, 0);
    p.ext_rec.clear(); p.ext_pos.clear(); p.ext_w.clear();
    for (const R &e : rs) {
        p.ext_tile[e.t + 1]++;
        p.ext_rec.push_back(e.r);
        p.ext_pos.push_back((e.loc % tz) | ((e.loc / tz) << 16));
        p.ext_w.insert(p.ext_w.end(), e.w, e.w + 16);
    }
    for (int t = 0; t < ntile; t++) p.ext_tile[t + 1] += p.ext_tile[t];
    p.ext_kind = kind;
}

// ---------------------------------------------------------------------------
// Batch: device memory of all shots of one call.

template <class T>
static size_t put(Layout &L, std::vector<std::pair<size_t, const void *>> &up,
                  std::vector<size_t> &upn, const std::vector<T> &v) {
    const size_t o = L.take(std::max<size_t>(v.size(), 1) * sizeof(T));
    if (!v.empty()) {
        up.push_back({o, v.data()});
        upn.push_back(v.size() * sizeof(T));
    }
    return o;
}

void Engine::build_batch(Batch &B, std::vector<ShotPlan> &plans, const BatchNeeds &need) {
    B.key.clear();   // the device tables no longer hold the plan any key named
    const int ns = (int)plans.size();
    Layout L;
    std::vector<std::pair<size_t, const void *>> up;
    std::vector<size_t> upn;
    struct Off {
        size_t u[kLevels], v2, vel, hz, hx, colsrc, taper;
        size_t inj_tile, inj_cell, inj_cptr, inj_coef, inj_tptr, inj_ttr, inj_traces, inj_tm;
        size_t injb_tile, injb_cell, injb_cptr, injb_coef, injb_tptr, injb_ttr;
        size_t inj_desc, inj_map, injb_desc, injb_map;
        size_t inj2[6], injb2[6];
        size_t ext_tile, ext_rec, ext_pos, ext_w, traces;
        size_t snap, snap_bytes, snap_scale, image, illum, born_m;
        size_t fs_row, fs_w;
        size_t sz, delp, rho, image2;
    };
    std::vector<Off> off(ns);
    const bool tb = tz == B2Z && tx == B2X;
    std::vector<RegionPlan> &rp = B.rp, &rpb = B.rpb;   // kept in the batch: uploads read them asynchronously
    rp.assign(tb ? ns : 0, RegionPlan());
    rpb.assign(tb && need.backward_injection ? ns : 0, RegionPlan());
    tb_ok = tb;
    int tile0 = 0;
    B.blk_shot_h.clear();
    B.colsrc_d.clear();
    for (int s = 0; s < ns; s++) {
        ShotPlan &p = plans[s];
        Off &o = off[s];
        const size_t nf = (size_t)p.LDZ * p.NX;
        for (int k = 0; k < kLevels; k++) o.u[k] = L.take(nf * sizeof(float));
        o.v2 = L.take(nf * sizeof(float));
        o.vel = L.take(nf * sizeof(float));
        o.sz = o.delp = o.rho = 0;
        if (has_rho()) {
            o.sz = L.take(nf * sizeof(float));
            o.delp = L.take(nf * sizeof(float));
            o.rho = L.take(nf * sizeof(float));
        }
        o.hz = L.take((size_t)16 * (p.NX - 8) * sizeof(float));
        o.hx = L.take((size_t)16 * (p.NZ - 8) * sizeof(float));
        o.colsrc = put(L, up, upn, p.colsrc);
        o.taper = put(L, up, upn, p.taper);
        o.fs_row = put(L, up, upn, p.fs_row);
        o.fs_w = put(L, up, upn, p.fs_w);
        o.inj_tile = put(L, up, upn, p.inj_tile);
        o.inj_cell = put(L, up, upn, p.inj_cell);
        o.inj_cptr = put(L, up, upn, p.inj_cptr);
        o.inj_coef = put(L, up, upn, p.inj_coef);
        o.inj_tptr = put(L, up, upn, p.inj_tptr);
        o.inj_ttr = put(L, up, upn, p.inj_ttr);
        inj_dense(p.ntz * p.ntx, tz * tx, p.inj_tile, p.inj_cell, p.inj_cptr, p.inj_tptr, p.inj_ttr, p.inj_desc,
                  p.inj_map);
        o.inj_desc = put(L, up, upn, p.inj_desc);
        o.inj_map = put(L, up, upn, p.inj_map);
        o.inj_traces = L.take((size_t)p.nt * std::max(p.n_inj_traces, 1) * sizeof(float));
        o.inj_tm = L.take((size_t)p.nt * std::max(p.n_inj_traces, 1) * sizeof(float));
        if (need.backward_injection) {
            o.injb_tile = put(L, up, upn, p.injb_tile);
            o.injb_cell = put(L, up, upn, p.injb_cell);
            o.injb_cptr = put(L, up, upn, p.injb_cptr);
            o.injb_coef = put(L, up, upn, p.injb_coef);
            o.injb_tptr = put(L, up, upn, p.injb_tptr);
            o.injb_ttr = put(L, up, upn, p.injb_ttr);
            inj_dense(p.ntz * p.ntx, tz * tx, p.injb_tile, p.injb_cell, p.injb_cptr, p.injb_tptr, p.injb_ttr,
                      p.injb_desc, p.injb_map);
            o.injb_desc = put(L, up, upn, p.injb_desc);
            o.injb_map = put(L, up, upn, p.injb_map);
        }
        auto put_region = [&](RegionPlan &r, size_t *o6) {
            o6[0] = put(L, up, upn, r.tile); o6[1] = put(L, up, upn, r.cell); o6[2] = put(L, up, upn, r.cptr);
            o6[3] = put(L, up, upn, r.coef); o6[4] = put(L, up, upn, r.tptr); o6[5] = put(L, up, upn, r.ttr);
            tb_ok = tb_ok && r.ok;
        };
        if (tb) {
            rp[s] = region_plan(p, p.inj_tile, p.inj_cell, p.inj_cptr, p.inj_coef, p.inj_tptr, p.inj_ttr);
            put_region(rp[s], o.inj2);
            if (need.backward_injection) {
                rpb[s] = region_plan(p, p.injb_tile, p.injb_cell, p.injb_cptr, p.injb_coef, p.injb_tptr, p.injb_ttr);
                put_region(rpb[s], o.injb2);
            }
        }
        o.ext_tile = put(L, up, upn, p.ext_tile);
        o.ext_rec = put(L, up, upn, p.ext_rec);
        o.ext_pos = put(L, up, upn, p.ext_pos);
        o.ext_w = put(L, up, upn, p.ext_w);
        o.traces = L.take((size_t)p.nt * p.ntr * sizeof(float));
        const size_t npix = (size_t)g.nz * p.nxw;
        o.snap_bytes = need.snapshots ? (size_t)std::max(p.ntsnap, 1) * npix * (need.snap_fp16 ? 2 : 4) : 16;
        o.snap = L.take(o.snap_bytes);
        o.snap_scale = L.take(need.snap_fp16 ? (size_t)std::max(p.ntsnap, 1) * p.ntz * p.ntx * (tx / 4) * sizeof(float)
                                             : 16);
        o.image = L.take(need.imaging ? npix * sizeof(float) : 16);
        o.illum = L.take(need.imaging ? npix * sizeof(float) : 16);
        o.image2 = L.take(need.rhovp ? npix * sizeof(float) : 16);   // zeroed with the illumination's range
        o.born_m = L.take(need.born ? npix * sizeof(float) : 16);
    }
    // tiles shot by shot, longest record first (stable in shot order)
    B.order.resize(ns);
    for (int s = 0; s < ns; s++) B.order[s] = s;
    std::stable_sort(B.order.begin(), B.order.end(), [&](int a, int b) { return plans[a].nt > plans[b].nt; });
    B.nt_sorted.assign(ns, 0);
    B.tiles_upto.assign(ns + 1, 0);
    for (int k = 0; k < ns; k++) {
        ShotPlan &p = plans[B.order[k]];
        p.tile0 = tile0;
        const int nt_ = p.ntz * p.ntx;
        for (int q = 0; q < nt_; q++) B.blk_shot_h.push_back(B.order[k]);
        tile0 += nt_;
        B.nt_sorted[k] = p.nt;
        B.tiles_upto[k + 1] = tile0;
    }
    B.nblocks = tile0;
    const size_t o_blk = put(L, up, upn, B.blk_shot_h);
    const size_t o_order = put(L, up, upn, B.order);
    const size_t o_part = L.take(need.snap_global ? (size_t)kSnapParts * ns * sizeof(float) : 16);
    const size_t o_shots = L.take(sizeof(ShotDev) * ns);
    char *base;
    if (B.ext_base) {   // caller-provided memory (the allocation-free entry point)
        if (L.off > B.ext_cap)
            throw EngineError(SFD_ENOMEM, "the batch needs " + std::to_string(L.off) + " bytes of device memory, " +
                                              "more than the " + std::to_string(B.ext_cap) + " the scratch array holds");
        base = B.ext_base;
    } else {
        B.mem.reserve(L.off);
        base = (char *)B.mem.p;
    }
    // zero everything once (wavefields, accumulators, traces) except the
    // snapshots, then upload plans.  A pass that stores snapshots writes
    // every slot 1..ntsnap before any step reads it, so they need no fill
    // (on a 751x2301 single-shot RTM that fill was 13 GB).
    auto zero = [&](size_t o, size_t n) {
        if (B.ext_base) {   // a fill kernel, not a memset node, for callers that capture
            hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)std::min<size_t>((n / 16 + 255) / 256 + 1, 8192)),
                               dim3(256), 0, stream, base + o, n);
            HIPCHK(hipGetLastError());
        } else {
            HIPCHK(hipMemsetAsync(base + o, 0, n, stream));
        }
    };
    {
        size_t z0 = 0;
        for (int s = 0; s < ns; s++) {
            if (off[s].snap > z0) zero(z0, off[s].snap - z0);
            z0 = off[s].snap + off[s].snap_bytes;
        }
        if (L.off > z0) zero(z0, L.off - z0);
    }
    // what a later call on the same tables must clear again: the levels and
    // the accumulators (image, illumination, extracted traces)
    B.zero_ranges.clear();
    for (int s = 0; s < ns; s++) {
        const Off &o = off[s];
        const size_t nf = (size_t)plans[s].LDZ * plans[s].NX * sizeof(float);
        for (int k = 0; k < kLevels; k++) B.zero_ranges.push_back({o.u[k], nf});
        B.zero_ranges.push_back({o.image, o.illum - o.image});
        B.zero_ranges.push_back({o.illum, o.born_m - o.illum});
        B.zero_ranges.push_back({o.traces, o.snap - o.traces});
    }
    for (size_t i = 0; i < up.size(); i++)
        HIPCHK(hipMemcpyAsync(base + up[i].first, up[i].second, upn[i], hipMemcpyHostToDevice, stream));
    B.shots_h.assign(ns, ShotDev());
    B.injb.assign(ns, InjSet());
    for (int s = 0; s < ns; s++) {
        ShotPlan &p = plans[s];
        Off &o = off[s];
        ShotDev &S = B.shots_h[s];
        memset(&S, 0, sizeof S);
        for (int k = 0; k < kLevels; k++) S.u[k] = (float *)(base + o.u[k]);
        S.v2 = (const float *)(base + o.v2);
        S.vel = (const float *)(base + o.vel);
        S.hz = (const float *)(base + o.hz);
        S.hx = (const float *)(base + o.hx);
        S.NZ = p.NZ; S.NX = p.NX; S.LDZ = p.LDZ; S.zoff = p.zoff;
        S.zsh = p.zsh; S.xsh = p.xsh; S.ntz = p.ntz; S.ntx = p.ntx; S.tile0 = p.tile0; S.tz = tz;
        S.nz = g.nz; S.nxw = p.nxw; S.i0 = p.i0; S.iz0 = p.iz0; S.nt = p.nt; S.ntsnap = p.ntsnap;
        S.mirror = 2 * p.iz0 - 2;
        S.ext_ghost = g.rec_type != 0;
        S.no_top = g.surf_type > 0;
        S.hrow0 = g.surf_type > 0 ? p.iz0 : 4;   // Boundary_3d's nz0 = 1 with a free surface
        S.snap_mask = p.snap_mask;
        S.fs_row = p.fs_row.empty() ? nullptr : (const int *)(base + o.fs_row);
        S.fs_w = p.fs_w.empty() ? nullptr : (const float *)(base + o.fs_w);
        S.sz = has_rho() ? (float *)(base + o.sz) : nullptr;
        S.delp = has_rho() ? (const float *)(base + o.delp) : nullptr;
        S.rho = has_rho() ? (const float *)(base + o.rho) : nullptr;
        S.image2 = need.rhovp ? (float *)(base + o.image2) : nullptr;
        S.inj_tile = (const int *)(base + o.inj_tile);
        S.inj_cell = (const int *)(base + o.inj_cell);
        S.inj_cptr = (const int *)(base + o.inj_cptr);
        S.inj_coef = (const InjCoef *)(base + o.inj_coef);
        S.inj_tptr = (const int *)(base + o.inj_tptr);
        S.inj_ttr = (const int *)(base + o.inj_ttr);
        tile_range(p.inj_tile, S.inj_lo, S.inj_hi);
        S.inj_desc = (const int *)(base + o.inj_desc);
        S.inj_map = (const int *)(base + o.inj_map);
        S.inj_traces = (const float *)(base + o.inj_traces);
        S.inj_tm = (float *)(base + o.inj_tm);
        S.inj_ntr = std::max(p.n_inj_traces, 1);
        if (need.backward_injection) {
            InjSet &I = B.injb[s];
            I.tile = (const int *)(base + o.injb_tile);
            I.cell = (const int *)(base + o.injb_cell);
            I.cptr = (const int *)(base + o.injb_cptr);
            I.coef = (const InjCoef *)(base + o.injb_coef);
            I.tptr = (const int *)(base + o.injb_tptr);
            I.ttr = (const int *)(base + o.injb_ttr);
            tile_range(p.injb_tile, I.lo, I.hi);
            I.desc = (const int *)(base + o.injb_desc);
            I.map = (const int *)(base + o.injb_map);
            if (tb) {
                I.tile2 = (const int *)(base + o.injb2[0]); I.cell2 = (const int *)(base + o.injb2[1]);
                I.cptr2 = (const int *)(base + o.injb2[2]); I.coef2 = (const InjCoef *)(base + o.injb2[3]);
                I.tptr2 = (const int *)(base + o.injb2[4]); I.ttr2 = (const int *)(base + o.injb2[5]);
            }
        }
        if (tb) {
            S.inj2_tile = (const int *)(base + o.inj2[0]); S.inj2_cell = (const int *)(base + o.inj2[1]);
            S.inj2_cptr = (const int *)(base + o.inj2[2]); S.inj2_coef = (const InjCoef *)(base + o.inj2[3]);
            S.inj2_tptr = (const int *)(base + o.inj2[4]); S.inj2_ttr = (const int *)(base + o.inj2[5]);
        }
        S.ext_tile = p.ext_tile.empty() ? nullptr : (const int *)(base + o.ext_tile);
        S.ext_rec = (const int *)(base + o.ext_rec);
        S.ext_pos = (const int *)(base + o.ext_pos);
        S.ext_w = (const float *)(base + o.ext_w);
        S.traces = (float *)(base + o.traces);
        S.ext_kind = p.ext_kind;
        tile_range(p.ext_tile, S.ext_lo, S.ext_hi);
        S.snap = base + o.snap;
        S.snap_scale = need.snap_fp16 ? (float *)(base + o.snap_scale) : nullptr;
        S.nsb = p.ntz * p.ntx * (tx / 4);
        S.image = (float *)(base + o.image);
        S.illum = (float *)(base + o.illum);
        S.born_m = (const float *)(base + o.born_m);
        S.taper = (const float *)(base + o.taper);
        S.dt2snap = (g.dt * (float)g.snapi) * (g.dt * (float)g.snapi);
        B.colsrc_d.push_back((const int *)(base + o.colsrc));
    }
    B.blk_shot = (const int *)(base + o_blk);
    B.order_d = (const int *)(base + o_order);
    B.snap_part = (float *)(base + o_part);
    B.shots = (ShotDev *)(base + o_shots);
    B.ns = ns;
    upload_shots(B);
    B.shots_built = B.shots_h;
}

// Back to the state build_batch left (zeroed levels, image, illumination
// and extracted traces; the shot descriptors as built, e.g. the forward
// injection set) for a call whose plan the device tables already hold.
void Engine::reset_batch(Batch &B) {
    char *base = (char *)B.mem.p;
    for (const auto &r : B.zero_ranges) HIPCHK(hipMemsetAsync(base + r.first, 0, r.second, stream));
    B.shots_h = B.shots_built;
    upload_shots(B);
}

// Everything a plan depends on: the kind of call and its flags, the shots'
// geometry and record lengths, the model (version), tile and strip settings.
std::string Engine::plan_key(const char *kind, const sfd_shot *shots, int ns, int a, int b) const {
    std::string k(kind);
    auto add = [&](const void *p, size_t n) { k.append((const char *)p, n); };
    const int hdr[8] = {ns, a, b, tz, tx, strip_px, time_block, xcd_remap};
    add(hdr, sizeof hdr);
    add(&model_version, sizeof model_version);
    for (int s = 0; s < ns; s++) {
        const sfd_shot &sh = shots[s];
        const int n2[2] = {sh.nt, sh.ntr};
        add(n2, sizeof n2);
        add(&sh.sz, sizeof sh.sz);
        add(&sh.sx, sizeof sh.sx);
        add(sh.rz, sizeof(float) * sh.ntr);
        add(sh.rx, sizeof(float) * sh.ntr);
    }
    return k;
}

int Batch::active_shots(int n) const {
    // nt_sorted is non-increasing: the first place with nt <= n ends the prefix
    return (int)(std::lower_bound(nt_sorted.begin(), nt_sorted.end(), n, std::greater<int>()) - nt_sorted.begin());
}

void Engine::upload_shots(Batch &B) {
    HIPCHK(hipMemcpyAsync((void *)B.shots, B.shots_h.data(), sizeof(ShotDev) * B.ns, hipMemcpyHostToDevice, stream));
}

// Window + pad every shot's model and build its Higdon tables.
void Engine::prepare_models(Batch &B, std::vector<ShotPlan> &plans, const float *d_model) {
    for (int s = 0; s < B.ns; s++) {
        const ShotPlan &p = plans[s];
        ShotDev &S = B.shots_h[s];
        dim3 blk(256), grd((p.NZ + 255) / 256, p.NX);
        hipLaunchKernelGGL(window_pad_kernel, grd, blk, 0, stream, d_model, g.nz, B.colsrc_d[s], p.nxw, g.ntaper,
                           p.ntop, p.NZ, p.NX, p.LDZ, p.zoff, (float *)S.vel, (float *)S.v2);
        const int nk = std::max(p.NX, p.NZ);
        hipLaunchKernelGGL(higdon_tables_kernel, dim3((nk + 255) / 256), dim3(256), 0, stream, S.vel, p.NZ, p.NX,
                           p.LDZ, p.zoff, g.dt, g.dz, g.dx, (float *)S.hz, (float *)S.hx);
        if (has_rho()) {
            // padded rho, rho2 along z (in sz), 1/rho2 along x (delp), v^2 rho
            Sinc9 s9;
            sfd_host_mksinc(s9.w, 9, 0.5f);
            const size_t nf = (size_t)p.LDZ * p.NX;
            hipLaunchKernelGGL(rho_pad_kernel, grd, blk, 0, stream, (const float *)d_rho, g.nz, B.colsrc_d[s], p.nxw,
                               g.ntaper, p.ntop, p.NZ, p.NX, p.LDZ, p.zoff, (float *)S.rho);
            hipLaunchKernelGGL(rho_half_kernel, grd, blk, 0, stream, S.rho, S.sz, p.NZ, p.NX, p.LDZ, p.zoff, s9, 0);
            hipLaunchKernelGGL(rho_half_kernel, grd, blk, 0, stream, (const float *)S.sz, (float *)S.delp, p.NZ, p.NX,
                               p.LDZ, p.zoff, s9, 1);
            hipLaunchKernelGGL(v2_rho_kernel, dim3(2048), dim3(256), 0, stream, (float *)S.v2, S.rho, nf);
        }
    }
    HIPCHK(hipGetLastError());
}

void Engine::zero_wavefields(Batch &B, std::vector<ShotPlan> &plans) {
    for (int s = 0; s < B.ns; s++) {
        const size_t nf = (size_t)plans[s].LDZ * plans[s].NX * sizeof(float);
        for (int k = 0; k < kLevels; k++) HIPCHK(hipMemsetAsync(B.shots_h[s].u[k], 0, nf, stream));
    }
}

StepArgs Engine::base_args(const Batch &B) const {
    StepArgs A;
    memset(&A, 0, sizeof A);
    A.shots = B.shots;
    A.blk_shot = B.blk_shot;
    A.shot_order = B.order_d;
    A.snap_part = B.snap_part;
    A.nblocks = B.nblocks;
    A.xcd_remap = xcd_remap;
    A.snapi = g.snapi;
    // FDcoefs.f90:25-36 scaled as FDtypes.f90:367-387 (coefpower 2); with
    // variable density the staggered first-derivative coefficients
    // (FDcoefs.f90:12-23) over delta (coefpower 1, ReadParam.f90:49-53)
    const float c0 = -205.f / 72.f, c1 = 8.f / 5, c2 = -1.f / 5.f, c3 = 8.f / 315.f, c4 = -1.f / 560.f;
    const float dzi = 1.f / g.dz, dxi = 1.f / g.dx;
    const float dz2 = dzi * dzi, dx2 = dxi * dxi;
    const float cc[5] = {c0, c1, c2, c3, c4};
    const float cr[5] = {0.f, 1225.f / 1024.f, -(1225.f / (1024.f * 15.f)), 1225.f / (1024.f * 125.f),
                         -(1225.f / (1024.f * 1715.f))};
    for (int k = 0; k < 5; k++) {
        A.c[k] = has_rho() ? cr[k] * dzi : cc[k] * dz2;
        A.c[5 + k] = has_rho() ? cr[k] * dxi : cc[k] * dx2;
    }
    A.rdz = dzi;
    A.rdx = dxi;
    const float dt2 = g.dt * g.dt;
    A.c12dt2 = 12.f * dt2;
    A.div11 = 1.f / 11.f;
    return A;
}

typedef void (*StepFn)(StepArgs);

// Tiles are 64 (or 56) rows x TX columns, one wavefront per TZ x PX strip;
// 56-row tiles are the two-step kernel's tiling (a wavefront then spans a
// tile and its 4-row rings), with 32 columns and 4-column strips.
static int pick_px(int TZ, int TX) {
    if (TZ == 56) return TX == 32 ? 4 : 0;
    return TZ != 64 ? 0 : (TX == 32 ? 4 : (TX == 64 ? 16 : 0));
}
template <int PX, int MODE, int TZ>
static StepFn step_fn(bool adj, bool single) {
    return adj ? (single ? (StepFn)fd_step_kernel<PX, true, true, MODE, TZ, false>
                         : (StepFn)fd_step_kernel<PX, true, false, MODE, TZ, false>)
               : (single ? (StepFn)fd_step_kernel<PX, false, true, MODE, TZ, false>
                         : (StepFn)fd_step_kernel<PX, false, false, MODE, TZ, false>);
}
// Mode 3 (step tiles + imaging blocks) is launched for one shot only.
template <int PX, int TZ>
static StepFn step_fn_image(bool adj, bool single) {
    if (!single) return nullptr;
    return adj ? (StepFn)fd_step_kernel<PX, true, true, 3, TZ, false>
               : (StepFn)fd_step_kernel<PX, false, true, 3, TZ, false>;
}
// Variable density: 64 x 32 tiles, 4-column strips, the one stencil for both
// time directions (the drivers pass FD_2D_derivatives_acoustic_forward_grid
// to both RTM passes), no Born term.
static StepFn pick_rho_kernel(int TZ, int TX, bool single, int mode, int px_override) {
    if (TZ != 64 || TX != 32 || (px_override && px_override != 4)) return nullptr;
    switch (mode) {
    case 0: return single ? (StepFn)fd_step_kernel<4, false, true, 0, 64, true> : (StepFn)fd_step_kernel<4, false, false, 0, 64, true>;
    case 1: return single ? (StepFn)fd_step_kernel<4, false, true, 1, 64, true> : (StepFn)fd_step_kernel<4, false, false, 1, 64, true>;
    case 3: return single ? (StepFn)fd_step_kernel<4, false, true, 3, 64, true> : nullptr;
    default: return nullptr;
    }
}
static StepFn pick_kernel(int TZ, int TX, bool adj, bool single, int mode, int px_override = 0, bool rho = false) {
    if (rho) return pick_rho_kernel(TZ, TX, single, mode, px_override);
    const int want = px_override ? px_override : pick_px(TZ, TX);
#define K(px, tz)                                                                                        \
    if (TZ == tz && TX == StepShape<px>::TX && want == px)                                               \
        return mode == 3   ? step_fn_image<px, tz>(adj, single)                                          \
               : mode == 2 ? step_fn<px, 2, tz>(adj, single)                                             \
                           : (mode == 1 ? step_fn<px, 1, tz>(adj, single) : step_fn<px, 0, tz>(adj, single));
    K(4, 64) K(8, 64) K(16, 64) K(4, 56)
#undef K
    return nullptr;
}

bool Engine::tile_supported(int TZ, int TX) { return pick_kernel(TZ, TX, false, false, 0) != nullptr; }

// The code the launch needs (fd_step_body MODE): 0 unless the call stores
// snapshots or images and some shot of the batch is at a step it = its
// (its % snapi == 0) of its record in launch n (its = n+1 forward, nt-n
// backward); then 2 when it images or adds the Born source, else 1.  Imaging
// (RTM, LSRTM, FWI) launches mode 3 (the lean step plus imaging blocks).
static int snap_launch(const Batch &B, const StepArgs &A) {
    if (!A.store_snap && A.imaging == IMG_NONE) return 0;
    const int na = B.active_shots(A.n);
    for (int k = 0; k < na; k++) {
        const int its = A.dir > 0 ? A.n + 1 : B.nt_sorted[k] - A.n;
        if (its % A.snapi == 0) return A.imaging != IMG_NONE ? 2 : 1;
    }
    return 0;
}

void Engine::launch_step(const Batch &B, StepArgs A, bool adj) {
    if (B.ns == 1) {
        A.single = 1;
        A.s0 = B.shots_h[0];
    }
    const int mode = snap_launch(B, A);
    // the v-rho imaging condition: its own kernel on u2, before the surface
    // and the step change it; the step then runs lean
    const bool rhovp = mode == 2 && A.imaging == IMG_RHOVP;
    if (rhovp) {
        const int nact0 = B.active_shots(A.n);
        int nxw = 1;
        for (int k = 0; k < nact0; k++) nxw = std::max(nxw, B.shots_h[B.order[k]].nxw);
        hipLaunchKernelGGL(image_rhovp_kernel, dim3(nxw, (g.nz + 255) / 256, nact0), dim3(256), 0, stream, A,
                           1.f / g.dz, 1.f / g.dx);
    }
    // reverse-pass imaging: one shot, mode 3; a batch, mode 0 + image_kernel
    const bool image = mode == 2 && A.imaging != IMG_BORN && !rhovp;
    StepFn f = pick_kernel(tz, tx, adj, A.single != 0, image ? (A.single ? 3 : 0) : rhovp ? 0 : mode, strip_px,
                           has_rho());
    if (!f) throw EngineError(SFD_EUNSUPPORTED, has_rho() ? "variable density runs on 64x32 tiles with 4-column strips"
                                                          : "tile shape not compiled");
    // the shots still running at step n: a prefix of the tile layout
    const int nact = B.active_shots(A.n);
    A.nblocks = B.tiles_upto[nact];
    int grid = A.nblocks;
    if (A.xcd_remap) grid = ((grid + 7) / 8) * 8;
    A.step_grid = grid;
    const int px = strip_px ? strip_px : pick_px(tz, tx);
    const int threads = 64 * ((px >= 8 ? 4 * px : 32) / px);
    if (image) {
        if ((tz != 64 && tz != 56) || (tx & (tx - 1)))
            throw EngineError(SFD_EUNSUPPORTED, "imaging needs 64- or 56-row, 2^k-column tiles");
        int nxw = 1;
        for (int s = 0; s < B.ns; s++) nxw = std::max(nxw, B.shots_h[s].nxw);
        const int cols = A.single ? kImgColsSingle : kImgColsBatch;
        A.img_gx = (nxw + cols - 1) / cols;
        A.img_gy = (g.nz + 255) / 256;
        A.nshots = B.ns;
        if (A.single) {
            const int groups = A.img_gx * A.img_gy, per_block = threads / 256;
            grid += (groups + per_block - 1) / per_block;
        }
    }
    if (surface()) {
        // Propagator.f90:176-198: ExtractData reads u2 before
        // Boundary_set_free_surface_grid changes it, so the receivers are
        // read first, then the surface is set; the step kernel skips
        // extraction.  (Snapshots and imaging read only rows at and below the
        // surface, which the surface kernel does not change.)
        if (A.extract) {
            hipLaunchKernelGGL(extract_kernel, dim3(A.nblocks), dim3(256), 0, stream, A, tx);
            A.extract = 0;
        }
        int nxmax = 1;
        for (int k = 0; k < nact; k++) nxmax = std::max(nxmax, B.shots_h[B.order[k]].NX - 8);
        hipLaunchKernelGGL(free_surface_kernel, dim3((nxmax + 255) / 256, nact), dim3(256), 0, stream, A);
    }
    if (mode == 1 && A.store_snap && A.snap_fp16 && A.snap_global) {
        // one FP16 scale per snapshot: max |u2| over each window before the step stores it
        hipLaunchKernelGGL(snap_max_kernel, dim3(kSnapParts, nact), dim3(256), 0, stream, A);
        hipLaunchKernelGGL(snap_scale_kernel, dim3(nact), dim3(256), 0, stream, A);
    }
    hipLaunchKernelGGL(f, dim3(grid), dim3(threads), 0, stream, A);
    if (image && !A.single) {
        const dim3 ig(A.img_gx, A.img_gy, nact);
        if (A.imaging == IMG_FWI)
            hipLaunchKernelGGL(image_kernel<true>, ig, dim3(256), 0, stream, A);
        else
            hipLaunchKernelGGL(image_kernel<false>, ig, dim3(256), 0, stream, A);
    }
}

// The two-step kernel applies with 64x32 tiles, region plans within its trace
// table, and no Born source term.
bool Engine::use_step2(const StepArgs &A) const {
    return time_block == 2 && tb_ok && tz == B2Z && tx == B2X && A.imaging != IMG_BORN && !A.snap_global &&
           !ghosts_or_surface() && !has_rho();
}

void Engine::launch_step2(const Batch &B, StepArgs A, bool adj) {
    if (B.ns == 1) {
        A.single = 1;
        A.s0 = B.shots_h[0];
    }
    // both steps of the launch lie in the records of the same shots (run_loop)
    A.nblocks = B.active_tiles(A.n);
    int grid = A.nblocks;
    if (A.xcd_remap) grid = ((grid + 7) / 8) * 8;
    A.step_grid = grid;
    typedef void (*Fn)(StepArgs);
    const Fn f = adj ? (A.single ? (Fn)fd_step2_kernel<true, true> : (Fn)fd_step2_kernel<true, false>)
                     : (A.single ? (Fn)fd_step2_kernel<false, true> : (Fn)fd_step2_kernel<false, false>);
    hipLaunchKernelGGL(f, dim3(grid), dim3(kThreads2), 0, stream, A);
}

// The step kernel reads injected traces time-major (one row per step): copy
// each shot's traces (as uploaded or as the residual kernel wrote them).
void Engine::traces_time_major(const Batch &B) {
    for (int s = 0; s < B.ns; s++) {
        const ShotDev &S = B.shots_h[s];
        const dim3 grid((S.nt + 31) / 32, (S.inj_ntr + 31) / 32);
        hipLaunchKernelGGL(traces_tm_kernel, grid, dim3(32, 8), 0, stream, S.inj_traces, S.inj_tm, S.nt, S.inj_ntr);
    }
    HIPCHK(hipGetLastError());
}

// Ends a stream capture that an exception interrupts, so the stream is
// usable again and the partial graph is freed.
struct CaptureGuard {
    hipStream_t s = nullptr;
    bool on = false;
    void begin(hipStream_t st) {
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        s = st;
        on = true;
    }
    hipGraph_t end() {
        hipGraph_t gr = nullptr;
        on = false;
        HIPCHK(hipStreamEndCapture(s, &gr));
        return gr;
    }
    ~CaptureGuard() {
        if (!on) return;
        hipGraph_t gr = nullptr;
        (void)hipStreamEndCapture(s, &gr);
        if (gr) (void)hipGraphDestroy(gr);
    }
};

// Algorithmic HBM bytes of step n (two-step launch: steps n and n + 1) over
// the shots running then: 24 B per FD point-update (read U_n, U_{n-1},
// U_{n-2}, U_{n-3}, v^2; write U_{n+1}), plus on a shot's snapshot steps the
// snapshot write (forward) or the snapshot read(s) and the image and
// illumination read-modify-write (reverse-time imaging), per window cell.
double Engine::step_bytes(const Batch &B, const StepArgs &A, int n, bool two) const {
    double b = 0;
    const double sb = A.snap_fp16 ? 2.0 : 4.0;
    for (int m = n; m < n + (two ? 2 : 1); m++) {
        const int na = B.active_shots(m);
        for (int k = 0; k < na; k++) {
            const ShotDev &S = B.shots_h[B.order[k]];
            b += 24.0 * (double)(S.NZ - 8) * (double)(S.NX - 8);
            const int its = A.dir > 0 ? m + 1 : S.nt - m;
            if (its % A.snapi) continue;
            const double npix = (double)S.nz * (double)S.nxw;
            if (A.store_snap) b += (A.snap_global ? sb + 4.0 : sb) * npix;   // + the scale's read of u2
            if (A.imaging == IMG_BORN) b += (sb + 8.0) * npix;   // snapshot, reflectivity, velocity
            else if (A.imaging == IMG_RHOVP) b += (3.0 * sb + 36.0) * npix;   // + rho, u2, image2
            else if (A.imaging != IMG_NONE) b += ((A.imaging == IMG_FWI ? 3.0 : 1.0) * sb + 16.0) * npix;
        }
    }
    return b;
}

// Run steps 0 .. nsteps-1 of the lockstep time loop over a batch.  The shots
// still running at step n (nt > n) are a prefix of the tile layout, and every
// launch covers exactly their tiles.  Steps n < min nt (every shot running)
// go in blocks of K steps, K = graph_chunk rounded up to a multiple of
// lcm(kLevels, snapi) so that buffer rotation and snapshot phases repeat from
// block to block: the K step launches plus one base-advance launch are
// captured once into a HIP graph and replayed, the kernels reading the
// block's first step from device memory (d_nbase).  The remaining steps (the
// ragged tail, where shots finish one after another and launches shrink)
// are captured in chunks of at most K steps with their step index and grid
// fixed, each chunk one graph launch, captured while the GPU runs the
// previous ones.  HIP events on the engine stream bracket every block and
// chunk: kernel_stats covers every step of the pass.
void Engine::run_loop(const Batch &B, StepArgs A, bool adj, int nsteps) {
    if (nsteps < 1) return;
    if (A.inject) traces_time_major(B);
    const int dir = A.dir > 0 ? 0 : 1;
    // the step kernel must exist before any capture starts
    if (!pick_kernel(tz, tx, adj, B.ns == 1, 0, strip_px, has_rho()))
        throw EngineError(SFD_EUNSUPPORTED, has_rho() ? "variable density runs on 64x32 tiles with 4-column strips"
                                                      : "tile shape not compiled");
    int period = kLevels;
    while (period % g.snapi) period += kLevels;
    const int K = std::max(period, (std::max(1, graph_chunk) + period - 1) / period * period);
    const int nfree = std::min(nsteps, B.nt_sorted.back());   // steps with every shot running
    const int nblk = nfree / K;
    const bool graphs_on = use_graphs && !eager_untimed;
    int *d_nb = nullptr;
    if (graphs_on && nblk > 0) {
        nbase.reserve(256);
        d_nb = (int *)nbase.p;
    }
    const bool two = use_step2(A);   // two steps per launch where both lie in the same shots' records
    auto one = [&](int n, const int *nb) {
        A.n = n;
        A.rot = n % kLevels;
        A.nbase = nb;
        launch_step(B, A, adj);
    };
    auto pair = [&](int n, const int *nb) {
        A.n = n;
        A.rot = n % kLevels;
        A.nbase = nb;
        launch_step2(B, A, adj);
    };
    const int inc = two ? 2 : 1;
    // the tail in chunks [n0, n1) of at most K steps
    std::vector<std::pair<int, int>> chunks;
    for (int n0 = nblk * K; n0 < nsteps; n0 += K) chunks.push_back({n0, std::min(nsteps, n0 + K)});
    const bool timed = event_every > 0 && !eager_untimed;
    const int nspan = nblk + (int)chunks.size();
    if (timed && (int)kev.size() < 2 * nspan) {
        const size_t old = kev.size();
        kev.resize(2 * nspan);
        for (size_t i = old; i < kev.size(); i++) HIPCHK(hipEventCreate(&kev[i]));
    }
    std::vector<double> span_bytes, span_pts;
    std::vector<int> span_steps;
    auto pts_of = [&](int n) {
        double p = 0;
        const int na = B.active_shots(n);
        for (int k = 0; k < na; k++) {
            const ShotDev &S = B.shots_h[B.order[k]];
            p += (double)(S.NZ - 8) * (double)(S.NX - 8);
        }
        return p;
    };
    std::vector<std::pair<hipGraph_t, hipGraphExec_t>> graphs;
    struct GraphsFree {
        std::vector<std::pair<hipGraph_t, hipGraphExec_t>> &g;
        ~GraphsFree() {
            for (auto &q : g) {
                if (q.second) (void)hipGraphExecDestroy(q.second);
                if (q.first) (void)hipGraphDestroy(q.first);
            }
        }
    } graphs_free{graphs};
    int span = 0;
    // ---- blocks of K steps with every shot running: one graph, replayed
    if (nblk > 0) {
        double bk = 0, pk = 0;
        for (int n = 0; n < K; n += inc) bk += step_bytes(B, A, n, two);
        pk = (double)K * pts_of(0);
        hipGraphExec_t exec = nullptr;
        if (graphs_on) {
            hipLaunchKernelGGL(set_base_kernel, dim3(1), dim3(1), 0, stream, d_nb, 0);
            CaptureGuard cg;
            cg.begin(stream);
            for (int n = 0; n < K; n += inc) two ? pair(n, d_nb) : one(n, d_nb);
            hipLaunchKernelGGL(advance_base_kernel, dim3(1), dim3(1), 0, stream, d_nb, K);
            hipGraph_t gr = cg.end();
            graphs.push_back({gr, nullptr});
            HIPCHK(hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0));
            graphs.back().second = exec;
        }
        for (int b = 0; b < nblk; b++, span++) {
            if (timed) HIPCHK(hipEventRecord(kev[2 * span], stream));
            if (exec) {
                HIPCHK(hipGraphLaunch(exec, stream));
            } else {
                for (int n = 0; n < K; n += inc) two ? pair(b * K + n, nullptr) : one(b * K + n, nullptr);
            }
            if (timed) HIPCHK(hipEventRecord(kev[2 * span + 1], stream));
            // the replayed block's bytes: its snapshot phases repeat every K steps
            span_bytes.push_back(bk);
            span_pts.push_back(pk);
            span_steps.push_back(K);
        }
    }
    // ---- the tail: chunks with fixed step indices and shrinking grids
    for (const auto &c : chunks) {
        double bk = 0, pk = 0;
        auto launches = [&]() {
            for (int n = c.first; n < c.second;) {
                const bool p2 = two && n + 1 < c.second && B.active_shots(n) == B.active_shots(n + 1);
                if (p2) {
                    pair(n, nullptr);
                    bk += step_bytes(B, A, n, true);
                    pk += pts_of(n) + pts_of(n + 1);
                    n += 2;
                } else {
                    one(n, nullptr);
                    bk += step_bytes(B, A, n, false);
                    pk += pts_of(n);
                    n++;
                }
            }
        };
        hipGraphExec_t exec = nullptr;
        if (graphs_on) {
            CaptureGuard cg;
            cg.begin(stream);
            launches();
            hipGraph_t gr = cg.end();
            graphs.push_back({gr, nullptr});
            HIPCHK(hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0));
            graphs.back().second = exec;
        }
        if (timed) HIPCHK(hipEventRecord(kev[2 * span], stream));
        if (exec) HIPCHK(hipGraphLaunch(exec, stream));
        else launches();
        if (timed) HIPCHK(hipEventRecord(kev[2 * span + 1], stream));
        span_bytes.push_back(bk);
        span_pts.push_back(pk);
        span_steps.push_back(c.second - c.first);
        span++;
    }
    HIPCHK(hipGetLastError());
    if (timed && span > 0) {
        HIPCHK(hipEventSynchronize(kev[2 * span - 1]));
        for (int b = 0; b < span; b++) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, kev[2 * b], kev[2 * b + 1]));
            kstat_us[dir] += 1000.0 * ms;
            kstat_n[dir] += span_steps[b];
            kstat_bytes[dir] += span_bytes[b];
            kstat_pts[dir] += span_pts[b];
        }
    }
    // the graphs must outlive their launches
    if (!graphs.empty()) HIPCHK(hipStreamSynchronize(stream));
}

void Engine::reset_kstats() {
    kstat_us[0] = kstat_us[1] = 0;
    kstat_n[0] = kstat_n[1] = 0;
    kstat_pts[0] = kstat_pts[1] = kstat_bytes[0] = kstat_bytes[1] = 0;
}

static int max_nt(const std::vector<ShotPlan> &plans) {
    int m = 0;
    for (const ShotPlan &p : plans) m = std::max(m, p.nt);
    return m;
}

double Engine::count_pts(const std::vector<ShotPlan> &plans) const {
    double pts = 0;
    for (const ShotPlan &p : plans) pts += (double)(p.NZ - 8) * (double)(p.NX - 8) * (double)p.nt;
    return pts;
}

// Upload the traces a shot injects (nt per trace, time fastest).
void Engine::upload_inj_traces(Batch &B, int s, const float *tr, size_t count) {
    HIPCHK(hipMemcpyAsync((void *)B.shots_h[s].inj_traces, tr, count * sizeof(float), hipMemcpyHostToDevice, stream));
}

void Engine::check_shots(const sfd_shot *shots, int nshots, int snaps, bool may_tune) {
    if (!have_model) throw EngineError(SFD_ESTATE, "no model: call sfd_set_model first");
    if (!shots || nshots < 1) throw EngineError(SFD_EINVAL, "no shots");
    for (int s = 0; s < nshots; s++)
        if (!shots[s].wavelet || shots[s].nt < 1 || shots[s].ntr < 1) throw EngineError(SFD_EINVAL, "bad shot descriptor");
    if (autotune_pending && may_tune) autotune(shots, nshots, snaps);
}

// ---- tile / strip autotuner -------------------------------------------------
// Plans the batch for every compiled step-kernel shape (64x32 tiles with 4- or
// 8-column wavefront strips, 64x64 tiles with 16-column strips, and 64x32
// tiles with two time steps per launch), times a short run of forward launches
// of each on the real batch with HIP events (per step), and keeps the fastest.
// When the call stores snapshots (snaps = its SFD_SNAP_* mode, -1 for none)
// the timed steps store them too, so the snapshot-step kernels, which carry
// the imaging code of the backward pass, are part of the measurement.  A
// candidate replaces the first (64x32, 4-column strips, one step) only when it
// is more than 2% faster.  Runs on the first call after sfd_set_tile(h, 0, 0).
void Engine::autotune(const sfd_shot *shots, int ns, int snaps) {
    struct Cand { int tz, tx, px, tb; };
    static const Cand cands[] = {{64, 32, 4, 1}, {64, 32, 8, 1}, {64, 64, 16, 1}, {B2Z, B2X, 4, 2}};
    int nmin = shots[0].nt;
    for (int s = 1; s < ns; s++) nmin = std::min(nmin, shots[s].nt);
    const int steps1 = std::max(1, std::min(32, nmin));         // one-step launches timed
    const int steps2 = std::min(32, nmin - 1) / 2 * 2;          // two-step: both steps inside every record
    double best = 1e30;
    Cand pick = cands[0];
    tune_log.clear();
    for (const Cand &cd : cands) {
        if (cd.tb == 2 && (steps2 < 2 || ghosts_or_surface())) continue;
        if (has_rho() && !(cd.tz == 64 && cd.tx == 32 && cd.px == 4 && cd.tb == 1)) continue;
        const int steps = cd.tb == 2 ? steps2 : steps1;
        tz = cd.tz; tx = cd.tx; strip_px = cd.px; time_block = cd.tb;
        std::vector<ShotPlan> plans(ns);
        for_shots(ns, [&](int s) {
            plan_shot(shots[s], plans[s]);
            plan_injection(plans[s], vel_h, 1, &shots[s].sz, &shots[s].sx, 1);
            plans[s].n_inj_traces = 1;
        });
        BatchNeeds need;
        need.snapshots = snaps >= 0;
        need.snap_fp16 = snaps == SFD_SNAP_FP16S || snaps == SFD_SNAP_FP16G;
        need.snap_global = snaps == SFD_SNAP_FP16G;
        build_batch(fwd_batch, plans, need);
        for (int s = 0; s < ns; s++) upload_inj_traces(fwd_batch, s, shots[s].wavelet, shots[s].nt);
        traces_time_major(fwd_batch);
        prepare_models(fwd_batch, plans, d_vel);
        StepArgs A = base_args(fwd_batch);
        A.dir = 1;
        A.inject = 1;
        A.store_snap = need.snapshots;
        A.snap_fp16 = need.snap_fp16;
        A.snap_global = need.snap_global;
        const bool two = use_step2(A);
        auto adv = [&](int n) {
            A.n = n; A.rot = n % kLevels;
            if (two) launch_step2(fwd_batch, A, false);
            else launch_step(fwd_batch, A, false);
        };
        const int inc = two ? 2 : 1;
        for (int n = 0; n + inc <= 4 && n + inc < nmin; n += inc) adv(n);   // warm-up
        HIPCHK(hipEventRecord(ev0, stream));
        for (int n = 0; n + inc <= steps; n += inc) adv(n);
        HIPCHK(hipEventRecord(ev1, stream));
        HIPCHK(hipEventSynchronize(ev1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
        const double us = 1000.0 * ms / steps;
        char buf[96];
        snprintf(buf, sizeof buf, "%s%dx%d/px%d/tb%d %.2f us", tune_log.empty() ? "" : "; ", cd.tz, cd.tx, cd.px,
                 cd.tb, us);
        tune_log += buf;
        const double score = &cd == &cands[0] ? 0.98 * us : us;
        if (score < best) {
            best = score;
            pick = cd;
        }
    }
    tz = pick.tz; tx = pick.tx; strip_px = pick.px; time_block = pick.tb;
    autotune_pending = false;
}

// ---- HBM capacity: shot waves --------------------------------------------
// Device bytes one shot adds to a batch (build_batch's layout: wavefield
// levels, model copies, injection / extraction tables, traces, snapshots and
// image buffers), from its window alone, with a margin for the plan tables.
size_t Engine::shot_bytes(const sfd_shot &s, const BatchNeeds &need) const {
    float oxw;
    int nxw;
    shot_window(g, s, &oxw, &nxw);
    const size_t NZ = (size_t)g.nz + top_rows(g) + g.ntaper + 8, NX = (size_t)nxw + 2 * g.ntaper + 8;
    const size_t LDZ = (NZ + 64 + 31) / 32 * 32;
    const size_t ntile = (NZ + tz) / tz * ((NX + tx) / tx + 1);
    const size_t npix = (size_t)g.nz * nxw, ntr = (size_t)std::max(s.ntr, 1), nt = (size_t)s.nt;
    const size_t ntsnap = (size_t)std::max(s.nt / g.snapi, 1);
    size_t b = (kLevels + 2 + (has_rho() ? 3 : 0)) * LDZ * NX * 4;   // levels, v^2, vel (+ set-up scratch, 1/rho2, rho)
    b += 16 * (NX + NZ) * 4 + 2 * NX * 4;                     // Higdon tables, window columns, taper
    b += 3 * nt * ntr * 4;                                    // injected (two layouts) and extracted traces
    b += ntr * 49 * 16 * 4 + ntr * 16 * 4 + ntile * 64;       // contributions (both sets, two-step regions)
    b += (ntr + 4) * 2 * 64 * 32 * 4;                         // dense cell maps of the injecting tiles
    if (need.snapshots) b += ntsnap * npix * (need.snap_fp16 ? 2 : 4) + (need.snap_fp16 ? ntsnap * ntile * 16 * 4 : 0);
    if (need.imaging) b += 2 * npix * 4;
    if (need.rhovp) b += npix * 4;
    if (need.born) b += npix * 4;
    return b + (b >> 4) + (1 << 20);
}

// Contiguous ranges of shots [first, end) whose batches fit the device: the
// memory free now plus the batch buffer already held (it is reused or
// replaced), less a reserve for the call's other buffers, or the
// SFD_OPT_BATCH_BYTES cap when that is set.  One wave when everything fits.
// Waves follow shot order, so images, gradients and misfits summed over
// waves are added in the same order as in one batch.
std::vector<std::pair<int, int>> Engine::plan_waves(const sfd_shot *shots, int ns, const BatchNeeds &need) {
    size_t avail;
    if (batch_bytes_cap > 0) {
        avail = batch_bytes_cap;
    } else {
        size_t fr = 0, tot = 0;
        HIPCHK(hipMemGetInfo(&fr, &tot));
        const size_t reserve = std::max<size_t>((size_t)2 << 30, tot / 50);
        avail = fr + fwd_batch.mem.cap;
        avail = avail > reserve ? avail - reserve : 0;
    }
    std::vector<std::pair<int, int>> w;
    size_t acc = 0;
    int first = 0;
    for (int s = 0; s < ns; s++) {
        const size_t b = shot_bytes(shots[s], need);
        if (s > first && acc + b > avail) {
            w.push_back({first, s});
            first = s;
            acc = 0;
        }
        acc += b;
        if (acc > avail && s == first && batch_bytes_cap == 0)
            throw EngineError(SFD_ENOMEM, "one shot needs " + std::to_string(b >> 20) + " MiB of device memory, " +
                                              std::to_string(avail >> 20) + " MiB are available");
    }
    w.push_back({first, ns});
    waves_last = (int)w.size();
    return w;
}

// ---- forward modeling (Acoustic_modeling_sep.f90 / AMOD) -------------------
// ext: a batch in caller-provided device memory (the allocation-free entry
// point): one wave, no autotune, eager untimed launches.
void Engine::forward(const sfd_shot *shots, int ns, float *traces_out, int extract_kind, Batch *ext) {
    check_shots(shots, ns, -1, ext == nullptr);
    reset_kstats();
    if (extract_kind < 0 || extract_kind > 3) throw EngineError(SFD_EINVAL, "bad extract kind");
    BatchNeeds need;
    const auto waves = ext ? std::vector<std::pair<int, int>>{{0, ns}} : plan_waves(shots, ns, need);
    Batch &fwd_batch = ext ? *ext : this->fwd_batch;
    struct EagerScope {   // no graphs, no events, no device step base during an allocation-free call
        Engine &e;
        bool old;
        EagerScope(Engine &e_, bool on) : e(e_), old(e_.eager_untimed) { e.eager_untimed = old || on; }
        ~EagerScope() { e.eager_untimed = old; }
    } eager_scope(*this, ext != nullptr);
    double setup_ms = 0, run_ms = 0;
    size_t off = 0;
    pts = 0;
    for (const auto &wv : waves) {
        const int s0 = wv.first, nw = wv.second - wv.first;
        auto ta = std::chrono::steady_clock::now();
        std::vector<ShotPlan> plans(nw);
        for_shots(nw, [&](int s) {
            const sfd_shot &sh = shots[s0 + s];
            plan_shot(sh, plans[s]);
            plan_injection(plans[s], vel_h, 1, &sh.sz, &sh.sx, 1);
            plans[s].n_inj_traces = 1;
            plan_extraction(plans[s], vel_h, sh, extract_kind);
        });
        build_batch(fwd_batch, plans, need);
        for (int s = 0; s < nw; s++)
            upload_inj_traces(fwd_batch, s, source_wavelet(shots[s0 + s], s), shots[s0 + s].nt);
        HIPCHK(hipStreamSynchronize(stream));   // the shifted copies are rewritten by the next wave
        prepare_models(fwd_batch, plans, d_vel);
        StepArgs A = base_args(fwd_batch);
        A.dir = 1;
        A.inject = 1;
        A.extract = 1;
        setup_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ta).count();
        HIPCHK(hipEventRecord(ev0, stream));
        run_loop(fwd_batch, A, false, max_nt(plans));
        HIPCHK(hipEventRecord(ev1, stream));
        HIPCHK(hipEventSynchronize(ev1));
        for (int s = 0; s < nw; s++) {
            const size_t n = (size_t)plans[s].nt * plans[s].ntr;
            HIPCHK(hipMemcpy(traces_out + off, fwd_batch.shots_h[s].traces, n * sizeof(float), hipMemcpyDeviceToHost));
            off += n;
        }
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
        run_ms += ms;
        pts += count_pts(plans);
    }
    timing[0] = run_ms; timing[1] = run_ms; timing[2] = 0;
    timing[3] = setup_ms;
}

// Add every shot's window image into device full-model buffers (mod_copy_image,
// nearest index, shots in order); the buffers are zeroed by the caller.
void Engine::accumulate_images(Batch &B, std::vector<ShotPlan> &plans, float *d_img, float *d_ill, float *d_img2) {
    std::vector<std::vector<std::pair<int, int>>> inv(g.nx);
    for (int s = 0; s < B.ns; s++)
        for (int j = 0; j < plans[s].nxw; j++) {
            const int ix = plans[s].colmap_raw[j];
            if (ix < 1 || ix > g.nx) continue;   // ExtractPadModel.f90:686
            inv[ix - 1].push_back({s, j});
        }
    std::vector<int> ptr(g.nx + 1, 0), ishot, icol;
    for (int ix = 0; ix < g.nx; ix++) {
        ptr[ix + 1] = ptr[ix] + (int)inv[ix].size();
        for (const auto &e : inv[ix]) {
            ishot.push_back(e.first);
            icol.push_back(e.second);
        }
    }
    Layout L;
    const size_t o_ptr = L.take(ptr.size() * 4), o_shot = L.take(std::max<size_t>(ishot.size(), 1) * 4),
                 o_col = L.take(std::max<size_t>(icol.size(), 1) * 4);
    const size_t o_ci = L.take(sizeof(CopyImg) * B.ns), o_ci2 = L.take(sizeof(CopyImg) * B.ns),
                 o_ci3 = L.take(sizeof(CopyImg) * B.ns);
    tmp_mem.reserve(L.off);
    char *base = (char *)tmp_mem.p;
    HIPCHK(hipMemcpyAsync(base + o_ptr, ptr.data(), ptr.size() * 4, hipMemcpyHostToDevice, stream));
    if (!ishot.empty()) {
        HIPCHK(hipMemcpyAsync(base + o_shot, ishot.data(), ishot.size() * 4, hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(base + o_col, icol.data(), icol.size() * 4, hipMemcpyHostToDevice, stream));
    }
    std::vector<CopyImg> ci(B.ns), ci2(B.ns), ci3(B.ns);
    for (int s = 0; s < B.ns; s++) {
        ci[s] = {B.shots_h[s].image, nullptr, plans[s].nxw};
        ci2[s] = {B.shots_h[s].illum, nullptr, plans[s].nxw};
        ci3[s] = {B.shots_h[s].image2, nullptr, plans[s].nxw};
    }
    HIPCHK(hipMemcpyAsync(base + o_ci, ci.data(), sizeof(CopyImg) * B.ns, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(base + o_ci2, ci2.data(), sizeof(CopyImg) * B.ns, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(base + o_ci3, ci3.data(), sizeof(CopyImg) * B.ns, hipMemcpyHostToDevice, stream));
    const dim3 grid((g.nz + 255) / 256, g.nx);
    if (d_img)
        hipLaunchKernelGGL(copy_image_kernel, grid, dim3(256), 0, stream, (const CopyImg *)(base + o_ci), g.nz, g.nx,
                           (const int *)(base + o_ptr), (const int *)(base + o_shot), (const int *)(base + o_col),
                           d_img);
    if (d_ill)
        hipLaunchKernelGGL(copy_image_kernel, grid, dim3(256), 0, stream, (const CopyImg *)(base + o_ci2), g.nz, g.nx,
                           (const int *)(base + o_ptr), (const int *)(base + o_shot), (const int *)(base + o_col),
                           d_ill);
    if (d_img2)
        hipLaunchKernelGGL(copy_image_kernel, grid, dim3(256), 0, stream, (const CopyImg *)(base + o_ci3), g.nz, g.nx,
                           (const int *)(base + o_ptr), (const int *)(base + o_shot), (const int *)(base + o_col),
                           d_img2);
    HIPCHK(hipGetLastError());
    // the host tables above must outlive their async transfers
    HIPCHK(hipStreamSynchronize(stream));
}

// ---- exact discrete adjoint of forward modeling (no reference counterpart) --
// sbar = F^T dbar per shot, F the map wavelet -> traces of Engine::forward
// with the given extraction.  Receivers become sources of the transposed
// extraction (injection plan kinds 1, 5, 3, 4 for SINC_V2, NEAREST,
// SINC_LSRTM, SINC_UNIT); three launches per step (adjoint_kernels.h).
// Exact transpose of born(): per shot the background pass is rerun with
// FP32 snapshots, then the K7 reverse sweep runs with the receivers' LSRTM
// sinc extraction transposed as its source, and k7_born_kernel collects the
// window's mbar at every snapshot step; the host adds the window into the
// model through the transpose of born()'s clamped column gather.
void Engine::born_adjoint_exact(const sfd_shot *shots, int ns, const float *dbar, float *mbar) {
    if (ghosts_or_surface()) throw EngineError(SFD_EUNSUPPORTED, "the exact adjoints have no free surface or ghosts");
    if (has_rho()) throw EngineError(SFD_EUNSUPPORTED, "variable density is built for forward modeling, RTM and sfd_afwi_gradient");
    check_shots(shots, ns, SFD_SNAP_FP32);
    if (!dbar || !mbar) throw EngineError(SFD_EINVAL, "NULL dbar/mbar");
    if (!(tz == 64 && (tx == 32 || tx == 64))) throw EngineError(SFD_EUNSUPPORTED, "tile shape not compiled");
    reset_kstats();
    std::fill(mbar, mbar + (size_t)g.nz * g.nx, 0.f);
    size_t off = 0;
    for (int s = 0; s < ns; s++) {
        const sfd_shot &sh = shots[s];
        std::vector<ShotPlan> plans(1);
        ShotPlan &p = plans[0];
        plan_shot(sh, p);
        plan_injection(p, vel_h, 1, &sh.sz, &sh.sx, 1);
        ShotPlan tmp = p;
        plan_injection(tmp, vel_h, sh.ntr, sh.rz, sh.rx, 3);   // transpose of the LSRTM sinc extraction
        p.take_backward(tmp);
        p.n_inj_traces = std::max(1, sh.ntr);
        BatchNeeds need;
        need.backward_injection = true;
        need.snapshots = true;
        Batch &B = fwd_batch;
        build_batch(B, plans, need);
        std::vector<float> w(sh.wavelet, sh.wavelet + sh.nt);
        second_difference(w, g.dt);
        upload_inj_traces(B, 0, w.data(), w.size());
        prepare_models(B, plans, d_vel);
        StepArgs A1 = base_args(B);
        A1.dir = 1;
        A1.inject = 1;
        A1.store_snap = 1;
        run_loop(B, A1, false, sh.nt);
        ShotDev S = B.shots_h[0];
        B.injb[0].apply(S);
        upload_inj_traces(B, 0, dbar + off, (size_t)sh.nt * sh.ntr);
        const size_t nf = (size_t)p.LDZ * p.NX, nw = (size_t)g.nz * p.nxw;
        Layout L;
        const size_t o_lev = L.take(6 * nf * 4), o_mb = L.take(nw * 4);
        tmp_mem.reserve(L.off);
        char *base = (char *)tmp_mem.p;
        HIPCHK(hipMemsetAsync(base + o_lev, 0, 6 * nf * 4, stream));
        HIPCHK(hipMemsetAsync(base + o_mb, 0, nw * 4, stream));
        const StepArgs SA = base_args(B);
        AdjStep A;
        memset(&A, 0, sizeof A);
        A.S = S;
        for (int q = 0; q < 10; q++) A.c[q] = SA.c[q];
        A.c12dt2 = SA.c12dt2;
        A.div11 = SA.div11;
        float *lev[5];
        for (int q = 0; q < 5; q++) lev[q] = (float *)(base + o_lev) + q * nf;
        A.W = (float *)(base + o_lev) + 5 * nf;
        float *d_mb = (float *)(base + o_mb);
        const int ntile = p.ntz * p.ntx;
        const unsigned ucells = (unsigned)(((size_t)p.NZ * p.NX + 255) / 256);
        const unsigned wcells = (unsigned)((nw + 255) / 256);
        for (int k = sh.nt; k >= 1; k--) {
            A.G = lev[0]; A.L0 = lev[1]; A.L1 = lev[2]; A.L2 = lev[3]; A.L3 = lev[4];
            A.k = k;
            if (tx == 32)
                hipLaunchKernelGGL((k7_gamma_kernel<64, 32>), dim3(ntile), dim3(256), 0, stream, A);
            else
                hipLaunchKernelGGL((k7_gamma_kernel<64, 64>), dim3(ntile), dim3(256), 0, stream, A);
            hipLaunchKernelGGL(k7_update_kernel, dim3(ucells), dim3(256), 0, stream, A);
            if (k % g.snapi == 0)
                hipLaunchKernelGGL(k7_born_kernel, dim3(wcells), dim3(256), 0, stream, A, k / g.snapi, d_mb);
            float *t = lev[0];
            lev[0] = lev[1]; lev[1] = lev[2]; lev[2] = lev[3]; lev[3] = lev[4]; lev[4] = t;
        }
        HIPCHK(hipGetLastError());
        std::vector<float> mw(nw);
        HIPCHK(hipMemcpyAsync(mw.data(), d_mb, nw * 4, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        for (int j = 0; j < p.nxw; j++)
            for (int i = 0; i < g.nz; i++) mbar[(size_t)p.colsrc[j] * g.nz + i] += mw[(size_t)j * g.nz + i];
        off += (size_t)sh.nt * sh.ntr;
    }
}
void Engine::adjoint_exact(const sfd_shot *shots, int ns, const float *dbar, float *sbar, int extract_kind) {
    if (ghosts_or_surface()) throw EngineError(SFD_EUNSUPPORTED, "the exact adjoints have no free surface or ghosts");
    if (has_rho()) throw EngineError(SFD_EUNSUPPORTED, "variable density is built for forward modeling, RTM and sfd_afwi_gradient");
    check_shots(shots, ns);
    if (!dbar || !sbar) throw EngineError(SFD_EINVAL, "NULL dbar/sbar");
    if (extract_kind < 0 || extract_kind > 3) throw EngineError(SFD_EINVAL, "bad extract_kind");
    if (!(tz == 64 && (tx == 32 || tx == 64))) throw EngineError(SFD_EUNSUPPORTED, "tile shape not compiled");
    static const int kind_of[4] = {1, 5, 3, 4};
    reset_kstats();
    size_t off = 0, sboff = 0;
    for (int s = 0; s < ns; s++) {
        const sfd_shot &sh = shots[s];
        std::vector<ShotPlan> plans(1);
        ShotPlan &p = plans[0];
        plan_shot(sh, p);
        plan_injection(p, vel_h, sh.ntr, sh.rz, sh.rx, kind_of[extract_kind]);
        p.n_inj_traces = std::max(1, sh.ntr);
        BatchNeeds need;
        Batch &B = fwd_batch;
        build_batch(B, plans, need);
        upload_inj_traces(B, 0, dbar + off, (size_t)sh.nt * sh.ntr);
        prepare_models(B, plans, d_vel);
        // the source footprint of Injection_sinc (j outer, i inner)
        int icz, icx;
        float dcz, dcx;
        find_i_d(&icz, &dcz, sh.sz, g.oz, g.dz);
        find_i_d(&icx, &dcx, sh.sx, p.oxw, g.dx);
        const int iz0 = icz - 1 + p.iz0, ix0 = icx - 1 + p.i0;
        const float v = padded_vel(vel_h, p, icz, icx), v2s = v * v;
        float wz[7], wx[7];
        sfd_host_mksinc(wz, 7, dcz / g.dz);
        sfd_host_mksinc(wx, 7, dcx / g.dx);
        std::vector<int> addr;
        std::vector<float> coef;
        for (int j = -3; j <= 3; j++)
            for (int i = -3; i <= 3; i++) {
                addr.push_back((ix0 + j) * p.LDZ + p.zoff + iz0 + i);
                coef.push_back(v2s * wz[3 + i] * wx[3 + j]);
            }
        const size_t nf = (size_t)p.LDZ * p.NX;
        Layout L;
        const size_t o_lev = L.take(6 * nf * 4), o_addr = L.take(addr.size() * 4), o_coef = L.take(coef.size() * 4),
                     o_sb = L.take((size_t)sh.nt * 4);
        tmp_mem.reserve(L.off);
        char *base = (char *)tmp_mem.p;
        HIPCHK(hipMemsetAsync(base + o_lev, 0, 6 * nf * 4, stream));
        HIPCHK(hipMemcpyAsync(base + o_addr, addr.data(), addr.size() * 4, hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(base + o_coef, coef.data(), coef.size() * 4, hipMemcpyHostToDevice, stream));
        const StepArgs SA = base_args(B);
        AdjStep A;
        memset(&A, 0, sizeof A);
        A.S = B.shots_h[0];
        A.nsrc = (int)addr.size();
        A.src_addr = (const int *)(base + o_addr);
        A.src_coef = (const float *)(base + o_coef);
        A.sbar = (float *)(base + o_sb);
        for (int q = 0; q < 10; q++) A.c[q] = SA.c[q];
        A.c12dt2 = SA.c12dt2;
        A.div11 = SA.div11;
        float *lev[5];
        for (int q = 0; q < 5; q++) lev[q] = (float *)(base + o_lev) + q * nf;
        A.W = (float *)(base + o_lev) + 5 * nf;
        const int ntile = p.ntz * p.ntx;
        const unsigned ucells = (unsigned)(((size_t)p.NZ * p.NX + 255) / 256);
        for (int k = sh.nt; k >= 1; k--) {
            A.G = lev[0]; A.L0 = lev[1]; A.L1 = lev[2]; A.L2 = lev[3]; A.L3 = lev[4];
            A.k = k;
            if (tx == 32)
                hipLaunchKernelGGL((k7_gamma_kernel<64, 32>), dim3(ntile), dim3(256), 0, stream, A);
            else
                hipLaunchKernelGGL((k7_gamma_kernel<64, 64>), dim3(ntile), dim3(256), 0, stream, A);
            hipLaunchKernelGGL(k7_update_kernel, dim3(ucells), dim3(256), 0, stream, A);
            hipLaunchKernelGGL(k7_source_kernel, dim3(1), dim3(64), 0, stream, A);
            float *tmp = lev[0];
            lev[0] = lev[1]; lev[1] = lev[2]; lev[2] = lev[3]; lev[3] = lev[4]; lev[4] = tmp;
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(sbar + sboff, base + o_sb, (size_t)sh.nt * 4,
                              hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        off += (size_t)sh.nt * sh.ntr;
        sboff += (size_t)sh.nt;
    }
}

// ---- RTM / LSRTM (to_disk_to_memory_RTM.f90:358-579) -------------------------
// Keep a device copy of traces (per shot (nt, ntr), concatenated in shot
// order) for RTM / LSRTM calls that pass no host traces.
void Engine::set_traces(const float *tr, size_t n) {
    resident_n = 0;
    if (!tr || n == 0) return;
    resident.reserve(n * sizeof(float));
    HIPCHK(hipMemcpy(resident.p, tr, n * sizeof(float), hipMemcpyHostToDevice));
    resident_n = n;
}

void Engine::rtm(const sfd_shot *shots, int ns, const float *traces_in, float *image, float *illum, int snap_mode,
                 bool lsrtm) {
    if (lsrtm && has_rho()) throw EngineError(SFD_EUNSUPPORTED, "variable density is built for forward modeling, RTM and sfd_afwi_gradient");
    check_shots(shots, ns, snap_mode);
    reset_kstats();
    if (!traces_in) {
        size_t total = 0;
        for (int s = 0; s < ns; s++) total += (size_t)shots[s].nt * shots[s].ntr;
        if (resident_n == 0) throw EngineError(SFD_EINVAL, "traces_in is NULL and no traces are resident");
        if (total != resident_n) throw EngineError(SFD_EINVAL, "resident traces do not match the shots' nt * ntr");
    }
    if (snap_mode != SFD_SNAP_FP32 && snap_mode != SFD_SNAP_FP16S && snap_mode != SFD_SNAP_FP16G)
        throw EngineError(SFD_EINVAL, "bad snap_mode");
    BatchNeeds need;
    need.backward_injection = true;
    need.snapshots = true;
    need.snap_fp16 = snap_mode == SFD_SNAP_FP16S || snap_mode == SFD_SNAP_FP16G;
    need.snap_global = snap_mode == SFD_SNAP_FP16G;
    need.imaging = true;
    Batch &B = fwd_batch;
    const auto waves = plan_waves(shots, ns, need);
    // full-model image and illumination summed over every wave on the device
    const size_t N = (size_t)g.nz * g.nx;
    acc_mem.reserve(2 * N * sizeof(float) + 256);
    float *d_img = (float *)acc_mem.p, *d_ill = d_img + N;
    HIPCHK(hipMemsetAsync(d_img, 0, 2 * N * sizeof(float), stream));
    double setup_ms = 0, fwd_ms = 0, bwd_ms = 0;
    size_t off = 0;
    pts = 0;
    for (const auto &wv : waves) {
        const int s0 = wv.first, nw = wv.second - wv.first;
        const sfd_shot *ws = shots + s0;
        auto ta = std::chrono::steady_clock::now();
        // a repeated one-wave call with the same plan (same shots, model, flags
        // and tile settings) reuses the tables already on the device
        const std::string key = waves.size() == 1 ? plan_key(lsrtm ? "lsrtm" : "rtm", ws, nw, snap_mode, 0) : "";
        if (plan_reuse && !key.empty() && !B.key.empty() && B.key == key) {
            reset_batch(B);
        } else {
            std::vector<ShotPlan> fresh(nw);
            for_shots(nw, [&](int s) {
                ShotPlan &p = fresh[s];
                plan_shot(ws[s], p);
                plan_injection(p, vel_h, 1, &ws[s].sz, &ws[s].sx, 1);
                // backward injection: receivers as sources (Injection_sinc / Injection_LSRTM_sinc)
                ShotPlan tmp = p;
                plan_injection(tmp, velb_h.empty() ? vel_h : velb_h, ws[s].ntr, ws[s].rz, ws[s].rx, lsrtm ? 3 : 1);
                p.take_backward(tmp);
                p.n_inj_traces = std::max(1, ws[s].ntr);
            });
            rtm_plans = std::move(fresh);
            build_batch(B, rtm_plans, need);
            B.key = key;
        }
        std::vector<ShotPlan> &plans = rtm_plans;
        // forward source (second-differenced for LSRTM, ReadSouVelRho.f90:318-326)
        for (int s = 0; s < nw; s++) {
            const float *w0 = source_wavelet(ws[s], s);
            std::vector<float> w(w0, w0 + ws[s].nt);
            if (lsrtm) second_difference(w, g.dt);
            upload_inj_traces(B, s, w.data(), w.size());
            HIPCHK(hipStreamSynchronize(stream));   // w is freed at the end of this iteration
        }
        prepare_models(B, plans, d_vel);
        setup_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ta).count();
        StepArgs A = base_args(B);
        A.dir = 1;
        A.inject = 1;
        A.store_snap = 1;
        A.snap_fp16 = need.snap_fp16;
        A.snap_global = need.snap_global;
        HIPCHK(hipEventRecord(ev0, stream));
        run_loop(B, A, false, max_nt(plans));
        HIPCHK(hipEventRecord(ev1, stream));
        // backward
        zero_wavefields(B, plans);
        if (d_velb) prepare_models(B, plans, d_velb);
        for (int s = 0; s < nw; s++) {
            ShotDev &S = B.shots_h[s];
            B.injb[s].apply(S);
            const size_t n = (size_t)ws[s].nt * ws[s].ntr;
            if (traces_in)
                upload_inj_traces(B, s, traces_in + off, n);
            else
                HIPCHK(hipMemcpyAsync((void *)S.inj_traces, (const float *)resident.p + off, n * sizeof(float),
                                      hipMemcpyDeviceToDevice, stream));
            off += n;
        }
        upload_shots(B);
        StepArgs Ab = base_args(B);
        Ab.dir = -1;
        Ab.inject = 1;
        Ab.snap_fp16 = need.snap_fp16;
        Ab.imaging = lsrtm ? IMG_LSRTM : IMG_RTM;
        run_loop(B, Ab, true, max_nt(plans));
        HIPCHK(hipEventRecord(ev2, stream));
        HIPCHK(hipEventSynchronize(ev2));
        float ms1 = 0, ms2 = 0;
        HIPCHK(hipEventElapsedTime(&ms1, ev0, ev1));
        HIPCHK(hipEventElapsedTime(&ms2, ev1, ev2));
        fwd_ms += ms1;
        bwd_ms += ms2;
        pts += 2 * count_pts(plans);
        accumulate_images(B, plans, image ? d_img : nullptr, illum ? d_ill : nullptr);
    }
    timing[0] = fwd_ms + bwd_ms; timing[1] = fwd_ms; timing[2] = bwd_ms;
    timing[3] = setup_ms;
    // image += sum over shots (the caller's arrays accumulate, as the reference's do)
    img_host.reserve(N * sizeof(float));
    float *h = (float *)img_host.p;
    for (int pass = 0; pass < 2; pass++) {
        float *dst = pass == 0 ? image : illum;
        if (!dst) continue;
        HIPCHK(hipMemcpyAsync(h, pass == 0 ? d_img : d_ill, N * sizeof(float), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        for (size_t q = 0; q < N; q++) dst[q] = dst[q] + h[q];
    }
}

// ---- Born modeling (to_disk_to_memory_BORN.f90:123-220) ----------------------
void Engine::born(const sfd_shot *shots, int ns, const float *refl, float *traces_out) {
    if (has_rho()) throw EngineError(SFD_EUNSUPPORTED, "variable density is built for forward modeling, RTM and sfd_afwi_gradient");
    check_shots(shots, ns, SFD_SNAP_FP32);
    reset_kstats();
    if (!refl || !traces_out) throw EngineError(SFD_EINVAL, "NULL refl/traces_out");
    BatchNeeds need;
    need.snapshots = true;
    need.born = true;
    Batch &B = fwd_batch;
    const auto waves = plan_waves(shots, ns, need);
    double setup_ms = 0, fwd_ms = 0, bwd_ms = 0;
    size_t off = 0;
    pts = 0;
    for (const auto &wv : waves) {
        const int s0 = wv.first, nw = wv.second - wv.first;
        const sfd_shot *ws = shots + s0;
        auto ta = std::chrono::steady_clock::now();
        std::vector<ShotPlan> plans(nw);
        for_shots(nw, [&](int s) {
            plan_shot(ws[s], plans[s]);
            plan_injection(plans[s], vel_h, 1, &ws[s].sz, &ws[s].sx, 1);
            plans[s].n_inj_traces = 1;
            plan_extraction(plans[s], vel_h, ws[s], EXT_SINC_LSRTM);
        });
        build_batch(B, plans, need);
        for (int s = 0; s < nw; s++) {
            std::vector<float> w(ws[s].wavelet, ws[s].wavelet + ws[s].nt);
            second_difference(w, g.dt);   // readsou with Born=.true.
            upload_inj_traces(B, s, w.data(), w.size());
            // reflectivity window (readref + mod_window_pad interior)
            std::vector<float> m((size_t)g.nz * plans[s].nxw);
            for (int j = 0; j < plans[s].nxw; j++)
                for (int i = 0; i < g.nz; i++) m[(size_t)j * g.nz + i] = refl[(size_t)plans[s].colsrc[j] * g.nz + i];
            HIPCHK(hipMemcpyAsync((void *)B.shots_h[s].born_m, m.data(), m.size() * 4, hipMemcpyHostToDevice, stream));
            HIPCHK(hipStreamSynchronize(stream));   // w and m are freed at the end of this iteration
        }
        prepare_models(B, plans, d_vel);
        setup_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ta).count();
        StepArgs A = base_args(B);
        A.dir = 1;
        A.inject = 1;
        A.store_snap = 1;
        HIPCHK(hipEventRecord(ev0, stream));
        run_loop(B, A, false, max_nt(plans));
        HIPCHK(hipEventRecord(ev1, stream));
        zero_wavefields(B, plans);
        StepArgs A2 = base_args(B);
        A2.dir = 1;
        A2.extract = 1;
        A2.imaging = IMG_BORN;
        run_loop(B, A2, false, max_nt(plans));
        HIPCHK(hipEventRecord(ev2, stream));
        HIPCHK(hipEventSynchronize(ev2));
        for (int s = 0; s < nw; s++) {
            const size_t n = (size_t)plans[s].nt * plans[s].ntr;
            HIPCHK(hipMemcpy(traces_out + off, B.shots_h[s].traces, n * sizeof(float), hipMemcpyDeviceToHost));
            off += n;
        }
        float ms1 = 0, ms2 = 0;
        HIPCHK(hipEventElapsedTime(&ms1, ev0, ev1));
        HIPCHK(hipEventElapsedTime(&ms2, ev1, ev2));
        fwd_ms += ms1;
        bwd_ms += ms2;
        pts += 2 * count_pts(plans);
    }
    timing[0] = fwd_ms + bwd_ms; timing[1] = fwd_ms; timing[2] = bwd_ms;
    timing[3] = setup_ms;
}

// ---- FWI misfit + gradient (compute_function_gradient.f90:208-350) ---------
// Shots run in waves (plan_waves); every wave adds its shots' per-trace misfits
// and window images into the call's buffers in shot order, then the
// post-processing chain runs once on the sums.
void Engine::fwi_gradient(const sfd_shot *shots, int ns, const float *dobs, const float *mute, const float *vpmask,
                          const sfd_fwi_opts *opt, double *misfit, float *grad, float *illum_out, float *resid_out) {
    if (has_rho()) throw EngineError(SFD_EUNSUPPORTED, "variable density is built for forward modeling, RTM and sfd_afwi_gradient");
    check_shots(shots, ns, opt ? opt->snap_mode : -1);
    reset_kstats();
    if (!opt || !misfit || !grad) throw EngineError(SFD_EINVAL, "NULL opt/misfit/grad");
    if (opt->snap_mode != SFD_SNAP_FP32 && opt->snap_mode != SFD_SNAP_FP16S && opt->snap_mode != SFD_SNAP_FP16G)
        throw EngineError(SFD_EINVAL, "bad snap_mode");
    if ((opt->do_z && opt->rect_z < 1) || (opt->do_x && opt->rect_x < 1)) throw EngineError(SFD_EINVAL, "bad smoothing length");
    if ((opt->do_z && opt->rect_z > g.nz) || (opt->do_x && opt->rect_x > g.nx))
        throw EngineError(SFD_EINVAL, "smoothing box longer than the model (Smoothing.f90:202 boxconv)");
    const int norm = opt->norm == 0 ? SFD_NORM_L2 : opt->norm;
    if (norm != SFD_NORM_L2 && norm != SFD_NORM_L1 && norm != SFD_NORM_HUBER && norm != SFD_NORM_CAUCHY &&
        norm != SFD_NORM_HYPERBOLIC)
        throw EngineError(SFD_EINVAL, "unknown norm (SFD_NORM_*)");
    if (!std::isfinite(opt->norm_thresh)) throw EngineError(SFD_EINVAL, "norm_thresh is not finite");
    long ndata = 0;
    int ntraces = 0;
    for (int s = 0; s < ns; s++) {
        ndata += (long)shots[s].nt * shots[s].ntr;
        ntraces += shots[s].ntr;
    }
    if (!dobs) {
        if (resident_n == 0) throw EngineError(SFD_EINVAL, "dobs is NULL and no traces are resident");
        if ((size_t)ndata != resident_n) throw EngineError(SFD_EINVAL, "resident traces do not match the shots' nt * ntr");
    }
    auto t0 = std::chrono::steady_clock::now();
    BatchNeeds need;
    need.backward_injection = true;
    need.snapshots = true;
    need.snap_fp16 = opt->snap_mode == SFD_SNAP_FP16S || opt->snap_mode == SFD_SNAP_FP16G;
    need.snap_global = opt->snap_mode == SFD_SNAP_FP16G;
    need.imaging = true;
    Batch &B = fwd_batch;
    const auto waves = plan_waves(shots, ns, need);

    // call-level FWI buffers
    const size_t N = (size_t)g.nz * g.nx;
    const int rblocks = 1024;
    Layout L;
    const size_t o_dobs = L.take(dobs ? ndata * 4 : 16), o_mute = L.take(mute ? ndata * 4 : 16), o_res = L.take(ndata * 4);
    const size_t o_refs = L.take(sizeof(TraceRef) * ntraces), o_adj = L.take(sizeof(float *) * ns);
    const size_t o_ft = L.take(sizeof(double) * ntraces), o_tmax = L.take(sizeof(float) * ntraces),
                 o_thr = L.take(sizeof(float) * ntraces), o_tbeg = L.take(sizeof(int) * (ns + 1));
    const size_t o_grad = L.take(N * 4), o_illu = L.take(N * 4), o_tmp = L.take(N * 4);
    const size_t o_mask = L.take(vpmask ? N * 4 : 16);
    const size_t o_bmax = L.take(rblocks * 4), o_bsum = L.take(rblocks * 8);
    const int maxn = std::max(g.nz, g.nx);
    const int maxr = std::max(opt->rect_z, opt->rect_x) + 1;
    const int nlines = std::max(g.nz, g.nx);
    const size_t o_pp = L.take((size_t)(maxn + maxr) * nlines * 4), o_qq = L.take((size_t)(maxn + 2 * maxr) * nlines * 4);
    fwi_mem.reserve(L.off);
    char *fb = (char *)fwi_mem.p;
    const float *d_dobs = dobs ? (const float *)(fb + o_dobs) : (const float *)resident.p;
    if (dobs) HIPCHK(hipMemcpyAsync(fb + o_dobs, dobs, ndata * 4, hipMemcpyHostToDevice, stream));
    if (mute) HIPCHK(hipMemcpyAsync(fb + o_mute, mute, ndata * 4, hipMemcpyHostToDevice, stream));
    if (vpmask) HIPCHK(hipMemcpyAsync(fb + o_mask, vpmask, N * 4, hipMemcpyHostToDevice, stream));
    float *d_grad = (float *)(fb + o_grad), *d_illu = (float *)(fb + o_illu), *d_tmp = (float *)(fb + o_tmp);
    HIPCHK(hipMemsetAsync(d_grad, 0, N * 4, stream));
    HIPCHK(hipMemsetAsync(d_illu, 0, N * 4, stream));
    double setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    double fwd_ms = 0, bwd_ms = 0;
    pts = 0;
    long doff = 0;     // first sample of the wave's traces in the concatenated data
    int toff = 0;      // first trace of the wave
    for (const auto &wv : waves) {
        const int s0 = wv.first, nw = wv.second - wv.first;
        const sfd_shot *ws = shots + s0;
        auto ta = std::chrono::steady_clock::now();
        std::vector<ShotPlan> plans(nw);
        for_shots(nw, [&](int s) {
            ShotPlan &p = plans[s];
            plan_shot(ws[s], p);
            plan_injection(p, vel_h, 1, &ws[s].sz, &ws[s].sx, 1);        // Injection_sinc (AMOD)
            plan_extraction(p, vel_h, ws[s], EXT_NEAREST);                 // Extraction_array_simple_afwi
            ShotPlan tmp = p;
            plan_injection(tmp, vel_h, ws[s].ntr, ws[s].rz, ws[s].rx, 2);  // Injection_simple (AGRAD)
            p.take_backward(tmp);
            p.n_inj_traces = std::max(1, ws[s].ntr);
        });
        build_batch(B, plans, need);
        for (int s = 0; s < nw; s++) upload_inj_traces(B, s, ws[s].wavelet, ws[s].nt);
        prepare_models(B, plans, d_vel);
        std::vector<TraceRef> refs;
        int wtr = 0;
        {
            long o = doff;
            for (int s = 0; s < nw; s++)
                for (int r = 0; r < ws[s].ntr; r++) {
                    refs.push_back({s, r, ws[s].nt, o});
                    o += ws[s].nt;
                    wtr++;
                }
        }
        std::vector<float *> adjp(nw);
        for (int s = 0; s < nw; s++) adjp[s] = (float *)B.shots_h[s].inj_traces;
        std::vector<int> tbeg(nw + 1, 0);   // the wave's traces of shot s: [tbeg[s], tbeg[s+1])
        for (int s = 0; s < nw; s++) tbeg[s + 1] = tbeg[s] + ws[s].ntr;
        HIPCHK(hipMemcpyAsync(fb + o_refs, refs.data(), sizeof(TraceRef) * wtr, hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(fb + o_adj, adjp.data(), sizeof(float *) * nw, hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(fb + o_tbeg, tbeg.data(), sizeof(int) * (nw + 1), hipMemcpyHostToDevice, stream));
        setup_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ta).count();

        // AMOD: forward modeling with nearest extraction and snapshots
        StepArgs A = base_args(B);
        A.dir = 1;
        A.inject = 1;
        A.extract = 1;
        A.store_snap = 1;
        A.snap_fp16 = need.snap_fp16;
        A.snap_global = need.snap_global;
        HIPCHK(hipEventRecord(ev0, stream));
        run_loop(B, A, false, max_nt(plans));
        // residual, misfit per trace and adjoint source into the injection buffers
        const TraceRef *d_refs = (const TraceRef *)(fb + o_refs);
        float *const *d_adj = (float *const *)(fb + o_adj);
        float *d_tmax = (float *)(fb + o_tmax) + toff, *d_thr = (float *)(fb + o_thr) + toff;
        hipLaunchKernelGGL(residual_kernel, dim3(wtr), dim3(256), 0, stream, d_refs, B.shots, d_dobs,
                           mute ? (const float *)(fb + o_mute) : nullptr, (float *)(fb + o_res), d_adj, d_tmax);
        hipLaunchKernelGGL(norm_thresh_kernel, dim3((nw + 63) / 64), dim3(64), 0, stream, nw,
                           (const int *)(fb + o_tbeg), (const float *)d_tmax, norm, opt->norm_thresh, d_thr);
        hipLaunchKernelGGL(misfit_kernel, dim3((wtr + 63) / 64), dim3(64), 0, stream, wtr, d_refs,
                           (const float *)(fb + o_res), (const float *)d_thr, norm, (double *)(fb + o_ft) + toff);
        if (norm != SFD_NORM_L2)
            hipLaunchKernelGGL(adjsrc_kernel, dim3(wtr), dim3(256), 0, stream, d_refs, d_adj, (const float *)d_thr, norm);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(ev1, stream));
        // AGRAD: reverse time, forward stencil, simple injection, FWI imaging condition
        zero_wavefields(B, plans);
        for (int s = 0; s < nw; s++) {
            ShotDev &S = B.shots_h[s];
            B.injb[s].apply(S);
        }
        upload_shots(B);
        StepArgs Ab = base_args(B);
        Ab.dir = -1;
        Ab.inject = 1;
        Ab.snap_fp16 = need.snap_fp16;
        Ab.imaging = IMG_FWI;
        run_loop(B, Ab, false, max_nt(plans));
        HIPCHK(hipEventRecord(ev2, stream));
        // mod_copy_image into the full model (summed over waves in shot order);
        // it waits for the stream, so the wave's host tables may go
        accumulate_images(B, plans, d_grad, d_illu);
        float ms1 = 0, ms2 = 0;
        HIPCHK(hipEventElapsedTime(&ms1, ev0, ev1));
        HIPCHK(hipEventElapsedTime(&ms2, ev1, ev2));
        fwd_ms += ms1;
        bwd_ms += ms2;
        pts += 2 * count_pts(plans);
        for (int s = 0; s < nw; s++) doff += (long)ws[s].nt * ws[s].ntr;
        toff += wtr;
    }
    auto tp = std::chrono::steady_clock::now();

    // post-processing chain
    if (raw_illum_host)   // the summed illumination before illupow and normalisation
        HIPCHK(hipMemcpyAsync(raw_illum_host, d_illu, N * 4, hipMemcpyDeviceToHost, stream));
    std::vector<double> ft(ntraces);
    HIPCHK(hipMemcpyAsync(ft.data(), fb + o_ft, sizeof(double) * ntraces, hipMemcpyDeviceToHost, stream));
    double scaling = 0;
    for (int s = 0; s < ns; s++) scaling += 2.0 * (double)shots[s].nt * (double)shots[s].ntr;
    hipLaunchKernelGGL(fwi_scale_kernel, dim3(rblocks), dim3(256), 0, stream, d_grad, d_illu, N, opt->illupow,
                       (float)scaling, (float *)(fb + o_bmax), (double *)(fb + o_bsum));
    std::vector<float> bmax(rblocks);
    std::vector<double> bsum(rblocks);
    HIPCHK(hipMemcpyAsync(bmax.data(), fb + o_bmax, rblocks * 4, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(bsum.data(), fb + o_bsum, rblocks * 8, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    double f = 0;
    for (int k = 0; k < ntraces; k++) f = f + ft[k];
    f = f / scaling;
    float mx = bmax[0];
    double ss = 0;
    const int used = (int)std::min<size_t>(rblocks, (N + 255) / 256);
    for (int k = 0; k < used; k++) {
        mx = std::max(mx, bmax[k]);
        ss = ss + bsum[k];
    }
    const double rn = sqrt(ss / (double)N);
    if (!std::isfinite(f) || !std::isfinite(rn))
        throw EngineError(SFD_ENOTFINITE, "misfit, gradient or illumination is not finite");
    const float *d_mask = vpmask ? (const float *)(fb + o_mask) : nullptr;
    hipLaunchKernelGGL(fwi_normalize_kernel, dim3(rblocks), dim3(256), 0, stream, d_grad, d_illu, N, mx, rn, d_mask);
    float *pp = (float *)(fb + o_pp), *qq = (float *)(fb + o_qq);
    if (opt->do_x) {   // triangle along x for every depth row (Smoothing.f90:109-116)
        hipLaunchKernelGGL(triangle_kernel, dim3((g.nz + 127) / 128), dim3(128), 0, stream, d_grad, d_tmp, g.nz, g.nx,
                           (size_t)1, (size_t)g.nz, opt->rect_x, pp, qq);
        std::swap(d_grad, d_tmp);
    }
    if (opt->do_z) {   // then along z for every column (:118-125)
        hipLaunchKernelGGL(triangle_kernel, dim3((g.nx + 127) / 128), dim3(128), 0, stream, d_grad, d_tmp, g.nx, g.nz,
                           (size_t)g.nz, (size_t)1, opt->rect_z, pp, qq);
        std::swap(d_grad, d_tmp);
    }
    if (d_mask) hipLaunchKernelGGL(mask_kernel, dim3(rblocks), dim3(256), 0, stream, d_grad, N, d_mask);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(grad, d_grad, N * 4, hipMemcpyDeviceToHost, stream));
    if (illum_out) HIPCHK(hipMemcpyAsync(illum_out, d_illu, N * 4, hipMemcpyDeviceToHost, stream));
    if (resid_out) HIPCHK(hipMemcpyAsync(resid_out, fb + o_res, ndata * 4, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    *misfit = f;
    timing[0] = fwd_ms + bwd_ms; timing[1] = fwd_ms; timing[2] = bwd_ms;
    timing[3] = setup_ms + std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tp).count();
}

// ---- 3DAFWI one-shot gradient (to_disk_to_memory_AFWI.f90:229-415) ---------
// The batch machinery of fwi_gradient with 3DAFWI's plugins: sinc extraction
// of gain 1 (kind 3) in the forward pass, the residual (same kernels:
// Compute_OF_RES_3D is Compute_OF_RES_ADJ without the residual copy), and a
// reverse pass with the adjoint-style stencil, Injection_sinc of the residual
// and FWI imaging, or with a density the density operator and the v-rho
// imaging condition (image_rhovp_kernel) into a second image.
void Engine::afwi_gradient(const sfd_shot *shots, int ns, const float *dobs, const float *mute,
                           const sfd_afwi_opts *opt, double *misfit, float *grad, float *illum_out) {
    if (!opt || !misfit || !grad) throw EngineError(SFD_EINVAL, "NULL opt/misfit/grad");
    check_shots(shots, ns, opt->snap_mode);
    reset_kstats();
    if (opt->snap_mode != SFD_SNAP_FP32 && opt->snap_mode != SFD_SNAP_FP16S && opt->snap_mode != SFD_SNAP_FP16G)
        throw EngineError(SFD_EINVAL, "bad snap_mode");
    const int norm = opt->norm == 0 ? SFD_NORM_L2 : opt->norm;
    if (norm != SFD_NORM_L2 && norm != SFD_NORM_L1 && norm != SFD_NORM_HUBER && norm != SFD_NORM_CAUCHY &&
        norm != SFD_NORM_HYPERBOLIC)
        throw EngineError(SFD_EINVAL, "unknown norm (SFD_NORM_*)");
    if (!std::isfinite(opt->norm_thresh)) throw EngineError(SFD_EINVAL, "norm_thresh is not finite");
    if (opt->vprho_param != 0 && opt->vprho_param != 1) throw EngineError(SFD_EINVAL, "vprho_param must be 0 or 1");
    const bool rho = has_rho();
    if (rho)
        for (int s = 0; s < ns; s++) {
            float oxw;
            int nxw;
            shot_window(g, shots[s], &oxw, &nxw);
            if (nxw < 2)
                throw EngineError(SFD_EINVAL, "the v-rho imaging condition needs a window of two or more columns");
        }
    long ndata = 0;
    int ntraces = 0;
    for (int s = 0; s < ns; s++) {
        ndata += (long)shots[s].nt * shots[s].ntr;
        ntraces += shots[s].ntr;
    }
    if (!dobs) {
        if (resident_n == 0) throw EngineError(SFD_EINVAL, "dobs is NULL and no traces are resident");
        if ((size_t)ndata != resident_n) throw EngineError(SFD_EINVAL, "resident traces do not match the shots' nt * ntr");
    }
    auto t0 = std::chrono::steady_clock::now();
    BatchNeeds need;
    need.backward_injection = true;
    need.snapshots = true;
    need.snap_fp16 = opt->snap_mode == SFD_SNAP_FP16S || opt->snap_mode == SFD_SNAP_FP16G;
    need.snap_global = opt->snap_mode == SFD_SNAP_FP16G;
    need.imaging = true;
    need.rhovp = rho;
    Batch &B = fwd_batch;
    const auto waves = plan_waves(shots, ns, need);
    const size_t N = (size_t)g.nz * g.nx;
    const int np = rho ? 2 : 1;
    Layout L;
    const size_t o_dobs = L.take(dobs ? ndata * 4 : 16), o_mute = L.take(mute ? ndata * 4 : 16), o_res = L.take(ndata * 4);
    const size_t o_refs = L.take(sizeof(TraceRef) * ntraces), o_adj = L.take(sizeof(float *) * ns);
    const size_t o_ft = L.take(sizeof(double) * ntraces), o_tmax = L.take(sizeof(float) * ntraces),
                 o_thr = L.take(sizeof(float) * ntraces), o_tbeg = L.take(sizeof(int) * (ns + 1));
    const size_t o_grad = L.take(np * N * 4), o_illu = L.take(N * 4);
    fwi_mem.reserve(L.off);
    char *fb = (char *)fwi_mem.p;
    const float *d_dobs = dobs ? (const float *)(fb + o_dobs) : (const float *)resident.p;
    if (dobs) HIPCHK(hipMemcpyAsync(fb + o_dobs, dobs, ndata * 4, hipMemcpyHostToDevice, stream));
    if (mute) HIPCHK(hipMemcpyAsync(fb + o_mute, mute, ndata * 4, hipMemcpyHostToDevice, stream));
    float *d_grad = (float *)(fb + o_grad), *d_illu = (float *)(fb + o_illu);
    HIPCHK(hipMemsetAsync(d_grad, 0, np * N * 4, stream));
    HIPCHK(hipMemsetAsync(d_illu, 0, N * 4, stream));
    double setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    double fwd_ms = 0, bwd_ms = 0;
    pts = 0;
    long doff = 0;
    int toff = 0;
    for (const auto &wv : waves) {
        const int s0 = wv.first, nw = wv.second - wv.first;
        const sfd_shot *ws = shots + s0;
        auto ta = std::chrono::steady_clock::now();
        std::vector<ShotPlan> plans(nw);
        for_shots(nw, [&](int s) {
            ShotPlan &p = plans[s];
            plan_shot(ws[s], p);
            plan_injection(p, vel_h, 1, &ws[s].sz, &ws[s].sx, 1);         // Injection_sinc
            plan_extraction(p, vel_h, ws[s], EXT_SINC_UNIT);               // Extraction_array_sinc_afwi
            ShotPlan tmp = p;
            plan_injection(tmp, vel_h, ws[s].ntr, ws[s].rz, ws[s].rx, 1);   // Injection_sinc of the residual
            p.take_backward(tmp);
            p.n_inj_traces = std::max(1, ws[s].ntr);
        });
        build_batch(B, plans, need);
        for (int s = 0; s < nw; s++) upload_inj_traces(B, s, source_wavelet(ws[s], s), ws[s].nt);
        HIPCHK(hipStreamSynchronize(stream));   // the shifted wavelets are reused by the next wave
        prepare_models(B, plans, d_vel);
        std::vector<TraceRef> refs;
        int wtr = 0;
        {
            long o = doff;
            for (int s = 0; s < nw; s++)
                for (int r = 0; r < ws[s].ntr; r++) {
                    refs.push_back({s, r, ws[s].nt, o});
                    o += ws[s].nt;
                    wtr++;
                }
        }
        std::vector<float *> adjp(nw);
        for (int s = 0; s < nw; s++) adjp[s] = (float *)B.shots_h[s].inj_traces;
        std::vector<int> tbeg(nw + 1, 0);
        for (int s = 0; s < nw; s++) tbeg[s + 1] = tbeg[s] + ws[s].ntr;
        HIPCHK(hipMemcpyAsync(fb + o_refs, refs.data(), sizeof(TraceRef) * wtr, hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(fb + o_adj, adjp.data(), sizeof(float *) * nw, hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(fb + o_tbeg, tbeg.data(), sizeof(int) * (nw + 1), hipMemcpyHostToDevice, stream));
        setup_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ta).count();

        StepArgs A = base_args(B);
        A.dir = 1;
        A.inject = 1;
        A.extract = 1;
        A.store_snap = 1;
        A.snap_fp16 = need.snap_fp16;
        A.snap_global = need.snap_global;
        HIPCHK(hipEventRecord(ev0, stream));
        run_loop(B, A, false, max_nt(plans));
        const TraceRef *d_refs = (const TraceRef *)(fb + o_refs);
        float *const *d_adj = (float *const *)(fb + o_adj);
        float *d_tmax = (float *)(fb + o_tmax) + toff, *d_thr = (float *)(fb + o_thr) + toff;
        hipLaunchKernelGGL(residual_kernel, dim3(wtr), dim3(256), 0, stream, d_refs, B.shots, d_dobs,
                           mute ? (const float *)(fb + o_mute) : nullptr, (float *)(fb + o_res), d_adj, d_tmax);
        hipLaunchKernelGGL(norm_thresh_kernel, dim3((nw + 63) / 64), dim3(64), 0, stream, nw,
                           (const int *)(fb + o_tbeg), (const float *)d_tmax, norm, opt->norm_thresh, d_thr);
        hipLaunchKernelGGL(misfit_kernel, dim3((wtr + 63) / 64), dim3(64), 0, stream, wtr, d_refs,
                           (const float *)(fb + o_res), (const float *)d_thr, norm, (double *)(fb + o_ft) + toff);
        if (norm != SFD_NORM_L2)
            hipLaunchKernelGGL(adjsrc_kernel, dim3(wtr), dim3(256), 0, stream, d_refs, d_adj, (const float *)d_thr, norm);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(ev1, stream));
        zero_wavefields(B, plans);
        for (int s = 0; s < nw; s++) B.injb[s].apply(B.shots_h[s]);
        upload_shots(B);
        StepArgs Ab = base_args(B);
        Ab.dir = -1;
        Ab.inject = 1;
        Ab.snap_fp16 = need.snap_fp16;
        Ab.imaging = rho ? IMG_RHOVP : IMG_FWI;
        run_loop(B, Ab, true, max_nt(plans));
        HIPCHK(hipEventRecord(ev2, stream));
        accumulate_images(B, plans, d_grad, d_illu, rho ? d_grad + N : nullptr);
        float ms1 = 0, ms2 = 0;
        HIPCHK(hipEventElapsedTime(&ms1, ev0, ev1));
        HIPCHK(hipEventElapsedTime(&ms2, ev1, ev2));
        fwd_ms += ms1;
        bwd_ms += ms2;
        pts += 2 * count_pts(plans);
        for (int s = 0; s < nw; s++) doff += (long)ws[s].nt * ws[s].ntr;
        toff += wtr;
    }
    auto tp = std::chrono::steady_clock::now();
    std::vector<double> ft(ntraces);
    HIPCHK(hipMemcpyAsync(ft.data(), fb + o_ft, sizeof(double) * ntraces, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(grad, d_grad, np * N * 4, hipMemcpyDeviceToHost, stream));
    if (illum_out) HIPCHK(hipMemcpyAsync(illum_out, d_illu, N * 4, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    double f = 0;
    for (int k = 0; k < ntraces; k++) f = f + ft[k];
    if (!std::isfinite(f)) throw EngineError(SFD_ENOTFINITE, "misfit is not finite");
    if (rho && opt->vprho_param == 1)   // ExtractPadModel.f90:643-646 on the model's vel and rho
        for (size_t q = 0; q < N; q++) {
            const float v = vel_h[q], r = rho_h[q];
            grad[q] = grad[q] - (r / v) * grad[N + q];
            grad[N + q] = (1.f / v) * grad[N + q];
        }
    *misfit = f;
    timing[0] = fwd_ms + bwd_ms; timing[1] = fwd_ms; timing[2] = bwd_ms;
    timing[3] = setup_ms + std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tp).count();
}

// MuteParam_compute_mask (Mute.f90:50-98): the per-trace onsets and the two
// ramp tables in the reference's FP32 expressions (:178-186, :204-212), then
// the mask on the device.
void Engine::mute_mask(const sfd_shot *shots, int ns, float ot, const sfd_mute_opts &o, float *mask_out) {
    if (ns < 1 || !shots || !mask_out) throw EngineError(SFD_EINVAL, "no shots or no output");
    if (!(o.v0 > 0.f) || !(o.v1 > 0.f) || !(o.tramp >= 0.f) || !std::isfinite(o.t0) || !std::isfinite(o.t1) ||
        !std::isfinite(ot))
        throw EngineError(SFD_EINVAL, "mute: v0, v1 > 0, tramp >= 0 and finite times required");
    const float dt = g.dt, pi = 3.1415926535f;
    const float nr_f = o.tramp / dt + 1.5f;
    if (!(nr_f < 1e6f)) throw EngineError(SFD_EINVAL, "mute: tramp / dt too large");
    const int nramp = (int)nr_f;
    std::vector<float> rtop(nramp + 1), rbot(nramp + 1);
    for (int k = 0; k <= nramp; k++) {
        const float a = (float)k * pi / (float)(2 * nramp);
        const float sn = sinf(a), cs = cosf(a);
        rtop[k] = sn * sn;
        rbot[k] = cs * cs;
    }
    std::vector<TraceRef> refs;
    std::vector<int2> onset;
    long off = 0;
    for (int s = 0; s < ns; s++) {
        const sfd_shot &sh = shots[s];
        if (sh.nt < 1 || sh.ntr < 0 || (sh.ntr > 0 && !sh.rx)) throw EngineError(SFD_EINVAL, "bad shot");
        for (int r = 0; r < sh.ntr; r++) {
            const float hx = sh.rx[r] - sh.sx, hy = 0.f;
            auto onset_of = [&](float t_init, float v) {
                const float q = (hx + hy) / v;
                const float tm = sqrtf(t_init * t_init + q * q);
                const float x = (tm - ot) / dt + 1.5f;
                if (!(x > -2e9f && x < 2e9f)) throw EngineError(SFD_EINVAL, "mute onset out of range");
                return (int)x;   // conversion to INTEGER truncates
            };
            refs.push_back({s, r, sh.nt, off});
            onset.push_back(make_int2(onset_of(o.t0, o.v0), onset_of(o.t1, o.v1)));
            off += sh.nt;
        }
    }
    const int ntr = (int)refs.size();
    if (ntr == 0) return;
    Layout L;
    const size_t o_refs = L.take(sizeof(TraceRef) * ntr), o_on = L.take(sizeof(int2) * ntr),
                 o_rt = L.take(sizeof(float) * (nramp + 1)), o_rb = L.take(sizeof(float) * (nramp + 1)),
                 o_m = L.take(sizeof(float) * off);
    tmp_mem.reserve(L.off);
    char *b = (char *)tmp_mem.p;
    HIPCHK(hipMemcpyAsync(b + o_refs, refs.data(), sizeof(TraceRef) * ntr, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(b + o_on, onset.data(), sizeof(int2) * ntr, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(b + o_rt, rtop.data(), sizeof(float) * (nramp + 1), hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(b + o_rb, rbot.data(), sizeof(float) * (nramp + 1), hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(mute_kernel, dim3(ntr), dim3(256), 0, stream, (const TraceRef *)(b + o_refs),
                       (const int2 *)(b + o_on), nramp, (const float *)(b + o_rt), (const float *)(b + o_rb),
                       (float *)(b + o_m));
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(mask_out, b + o_m, sizeof(float) * off, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
}

// ---- Butterworth band-pass (Bandpass.f90:89-147 bandpass, bandpass.c) ----
// Section coefficients in the reference's own mix of float and double
// (bandpass.c:82-111 highcut_, :193-236 lowcut_), so the device recursion
// sees the reference's values; returns the number of sections.
static int bp_sections(int *np, int phase) {
    const int nodd = *np % 2;
    if (phase == 0) *np = nodd ? (*np + 1) / 2 : *np / 2;
    return (*np + 1) / 2;
}
static const double kBpPi = 3.14159265358979323846;

static int highcut_coefs(float fhi, int nphi, int phase, float *c) {
    int np = nphi;
    const int nsec = bp_sections(&np, phase), nodd = np % 2;
    if (nsec > kBpMaxSections) return -1;
    const float a = 2.0 * tan(kBpPi * fhi), aa = a * a, aap4 = aa + 4.0;
    const float dtheta = kBpPi / np, theta0 = nodd ? 0.0f : dtheta / 2.0;
    if (nodd) {
        c[0] = a / (a + 2.0);
        c[1] = c[0];
        c[2] = 0.0f;
        c[3] = (a - 2.0) / (a + 2.0);
        c[4] = 0.0f;
    }
    for (int j = nodd; j < nsec; j++) {
        const double cc = 4.0 * a * cos(theta0 + j * dtheta);
        float *b = c + 5 * j;
        b[0] = aa / (aap4 + cc);
        b[1] = 2.0 * b[0];
        b[2] = b[0];
        b[3] = (2.0 * aa - 8.0) / (aap4 + cc);
        b[4] = (aap4 - cc) / (aap4 + cc);
    }
    return nsec;
}

static int lowcut_coefs(float flo, int nplo, int phase, float *c) {
    int np = nplo;
    const int nsec = bp_sections(&np, phase), nodd = np % 2;
    if (nsec > kBpMaxSections) return -1;
    const float fno2 = 0.25f;
    const float a = 2. * sin(kBpPi * fno2) / cos(kBpPi * fno2), aa = a * a, aap4 = aa + 4;
    const float e = -cos(kBpPi * (flo + fno2)) / cos(kBpPi * (flo - fno2)), ee = e * e;
    const float dtheta = kBpPi / np, theta0 = nodd ? 0 : dtheta / 2;
    if (nodd) {
        const float b1 = a / (a + 2), b2 = (a - 2) / (a + 2), den = 1.0 - b2 * e;
        c[0] = b1 * (1. - e) / den;
        c[1] = -c[0];
        c[2] = 0.f;
        c[3] = (e - b2) / den;
        c[4] = 0.f;
    }
    for (int j = nodd; j < nsec; j++) {
        const double cc = 4. * a * cos(theta0 + j * dtheta);
        const float b1 = aa / (aap4 + cc), b2 = (2. * aa - 8.) / (aap4 + cc), b3 = (aap4 - cc) / (aap4 + cc);
        const float den = 1. - b2 * e + b3 * ee;
        float *d = c + 5 * j;
        d[0] = b1 * (1. - e) * (1. - e) / den;
        d[1] = -2. * d[0];
        d[2] = d[0];
        d[3] = (2. * e * (1. + b3) - b2 * (1. + ee)) / den;
        d[4] = (ee - b2 * e + b3) / den;
    }
    return nsec;
}

void Engine::bandpass(const sfd_bandpass_opts &o, int nt, int ntr, float *traces) {
    if (nt < 1 || ntr < 0 || (ntr > 0 && !traces)) throw EngineError(SFD_EINVAL, "bandpass: bad nt, ntr or traces");
    if (o.nplo < 0 || o.nphi < 0 || !std::isfinite(o.flo) || !std::isfinite(o.fhi))
        throw EngineError(SFD_EINVAL, "bandpass: negative pole count or non-finite cut-off");
    const float flo_t = o.flo * g.dt, fhi_t = o.fhi * g.dt;
    if (flo_t < 0.0001f && fhi_t > 0.4999f)   // Bandpass.f90:115 erexit
        throw EngineError(SFD_EINVAL, "bandpass: specify flo or fhi");
    std::vector<float> coef(2 * kBpMaxSections * 5, 0.f);
    int nlo = 0, nhi = 0;
    if (flo_t > 0.000001f) nlo = lowcut_coefs(flo_t, o.nplo, o.phase, coef.data());
    if (fhi_t < 4.999999f) nhi = highcut_coefs(fhi_t, o.nphi, o.phase, coef.data() + kBpMaxSections * 5);
    if (nlo < 0 || nhi < 0)
        throw EngineError(SFD_EUNSUPPORTED, "bandpass: more than " + std::to_string(kBpMaxSections) +
                                                " filter sections (poles > 16, or > 32 at zero phase)");
    if (ntr == 0) return;
    const float pi = 3.1415926535f;   // Bandpass.f90:100
    float tlo[kBpPad], thi[kBpPad];
    for (int j = 0; j < kBpPad; j++) {
        const float s1 = sinf((float)j * pi / (float)(2 * (kBpPad - 1)));
        const float s2 = sinf((float)(kBpPad - 1 - j) * pi / (float)(2 * (kBpPad - 1)));
        tlo[j] = s1 * s1;
        thi[j] = s2 * s2;
    }
    const int n1 = nt + 2 * kBpPad;
    const int ch = std::min(ntr, 1 << 16);
    const int ld = (ch + 63) & ~63;
    Layout L;
    const size_t o_c = L.take(coef.size() * 4), o_t = L.take(2 * kBpPad * 4), o_in = L.take((size_t)ch * nt * 4),
                 o_w = L.take((size_t)n1 * ld * 4);
    tmp_mem.reserve(L.off);
    char *b = (char *)tmp_mem.p;
    float *d_c = (float *)(b + o_c), *d_t = (float *)(b + o_t), *d_in = (float *)(b + o_in), *d_w = (float *)(b + o_w);
    HIPCHK(hipMemcpyAsync(d_c, coef.data(), coef.size() * 4, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(d_t, tlo, kBpPad * 4, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(d_t + kBpPad, thi, kBpPad * 4, hipMemcpyHostToDevice, stream));
    for (int k0 = 0; k0 < ntr; k0 += ch) {
        const int m = std::min(ch, ntr - k0);
        float *h = traces + (size_t)k0 * nt;
        HIPCHK(hipMemcpyAsync(d_in, h, (size_t)m * nt * 4, hipMemcpyHostToDevice, stream));
        const dim3 tg((nt + 63) / 64, (m + 63) / 64);
        hipLaunchKernelGGL(bp_pad_kernel, tg, dim3(256), 0, stream, (const float *)d_in, nt, m, ld,
                           (const float *)d_t, (const float *)(d_t + kBpPad), d_w);
        hipLaunchKernelGGL(bp_iir_kernel, dim3((m + 63) / 64), dim3(64), 0, stream, d_w, ld, m, n1,
                           (const float *)d_c, nlo, nhi, o.phase);
        hipLaunchKernelGGL(bp_unpad_kernel, tg, dim3(256), 0, stream, (const float *)d_w, nt, m, ld, d_in);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(h, d_in, (size_t)m * nt * 4, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }
}

void ShotPlan::take_backward(const ShotPlan &o) {
    injb_tile = o.inj_tile; injb_cell = o.inj_cell; injb_cptr = o.inj_cptr;
    injb_coef = o.inj_coef; injb_tptr = o.inj_tptr; injb_ttr = o.inj_ttr;
}

void second_difference(std::vector<float> &w, float dt) {
    // ReadSouVelRho.f90:318-326
    std::vector<float> o(w);
    const int nt = (int)w.size();
    for (int it = 0; it < nt; it++) w[it] = 0.f;
    for (int it = 2; it <= nt - 1; it++) w[it - 1] = w[it - 1] + (-2.f * o[it - 1] + o[it - 2] + o[it]) / (dt * dt);
}

// ---------------------------------------------------------------------------
// Allocation-free forward modeling.  Its batch (wavefield levels, model
// copies, plan tables, traces) lives in this file-scope device array instead
// of memory the engine allocates.  Each GPU the library runs on holds its own
// copy at its own address until the process ends; the host-side name below is
// a stand-in, not that address: host code uses only the per-device address
// hipGetSymbolAddress returns (taken once per device by sfd_create, with the
// device current).  Every call on a GPU shares that GPU's array, and the
// array keeps the last call's contents: a call zeroes what must start at zero
// before its first step, and calls that overlap on one GPU race.
static constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static std::mutex g_scratch_mu;
static std::map<int, char *> g_scratch_addr;   // device -> its copy of scratch

static char *scratch_address(int dev, std::string &err) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_addr.find(dev);
    if (it != g_scratch_addr.end()) return it->second;
    void *ptr = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (rc != hipSuccess || !ptr) {
        err = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(rc);
        return nullptr;
    }
    g_scratch_addr[dev] = (char *)ptr;
    return (char *)ptr;
}

// ---------------------------------------------------------------------------
// C-ABI

struct sfd_ctx {
    Engine *e = nullptr;
    std::string err;
};

static std::string g_create_err;

template <class F>
static int guarded(sfd_handle h, F &&f) {
    if (!h || !h->e) return SFD_EINVAL;
    try {
        f(*h->e);
        h->err.clear();
        return SFD_OK;
    } catch (const EngineError &e) {
        h->err = e.msg;
        return e.code;
    } catch (const std::exception &e) {
        h->err = e.what();
        return SFD_EINVAL;
    }
}

extern "C" {

int sfd_create(sfd_handle *h, const sfd_grid *g) {
    if (!h || !g) return SFD_EINVAL;
    *h = nullptr;
    try {
        std::unique_ptr<sfd_ctx> c(new sfd_ctx());
        c->e = new Engine(*g);
        // the allocation-free entry point's array, for the device current now
        int dev = 0;
        if (hipGetDevice(&dev) == hipSuccess) c->e->scratch_dev = scratch_address(dev, c->e->scratch_err);
        else c->e->scratch_err = "no current HIP device";
        *h = c.release();
        return SFD_OK;
    } catch (const EngineError &e) {
        g_create_err = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        g_create_err = "host allocation failed";
        return SFD_ENOMEM;
    } catch (const std::exception &e) {
        g_create_err = e.what();
        return SFD_EINVAL;
    }
}

int sfd_destroy(sfd_handle h) {
    if (!h) return SFD_EINVAL;
    delete h->e;
    delete h;
    return SFD_OK;
}

const char *sfd_last_error(sfd_handle h) {
    if (!h) return g_create_err.c_str();
    return h->err.c_str();
}

int sfd_device_info(char *name, int len, int *cu_count, size_t *hbm_bytes) {
    hipDeviceProp_t p;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) return SFD_EHIP;
    if (name && len > 0) {
        snprintf(name, (size_t)len, "%s%s(%s)", p.name, p.name[0] ? " " : "", p.gcnArchName);
    }
    if (cu_count) *cu_count = p.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = p.totalGlobalMem;
    return SFD_OK;
}

int sfd_set_model(sfd_handle h, const float *vel, const float *vel_bwd) {
    return guarded(h, [&](Engine &e) { e.set_model(vel, vel_bwd); });
}

int sfd_forward(sfd_handle h, const sfd_shot *shots, int nshots, float *traces_out, int extract_kind) {
    return guarded(h, [&](Engine &e) {
        if (!traces_out) throw EngineError(SFD_EINVAL, "traces_out is NULL");
        e.forward(shots, nshots, traces_out, extract_kind);
    });
}

int sfd_forward_noalloc(sfd_handle h, const sfd_shot *shots, int nshots, float *traces_out, int extract_kind) {
    return guarded(h, [&](Engine &e) {
        if (!traces_out) throw EngineError(SFD_EINVAL, "traces_out is NULL");
        if (!e.scratch_dev) throw EngineError(SFD_EHIP, "device scratch array unavailable: " + e.scratch_err);
        e.noalloc_batch.ext_base = e.scratch_dev;
        e.noalloc_batch.ext_cap = kScratchBytes;
        e.forward(shots, nshots, traces_out, extract_kind, &e.noalloc_batch);
    });
}

int sfd_rtm(sfd_handle h, const sfd_shot *shots, int nshots, const float *traces_in, float *image, float *illum,
            int snap_mode) {
    return guarded(h, [&](Engine &e) { e.rtm(shots, nshots, traces_in, image, illum, snap_mode, false); });
}

int sfd_set_traces(sfd_handle h, const float *traces, size_t count) {
    return guarded(h, [&](Engine &e) { e.set_traces(traces, count); });
}

int sfd_lsrtm(sfd_handle h, const sfd_shot *shots, int nshots, const float *traces_in, float *image, float *illum) {
    return guarded(h, [&](Engine &e) { e.rtm(shots, nshots, traces_in, image, illum, SFD_SNAP_FP32, true); });
}

int sfd_born(sfd_handle h, const sfd_shot *shots, int nshots, const float *refl, float *traces_out) {
    return guarded(h, [&](Engine &e) { e.born(shots, nshots, refl, traces_out); });
}

int sfd_fwi_gradient(sfd_handle h, const sfd_shot *shots, int nshots, const float *dobs, const float *mute,
                     const float *vpmask, const sfd_fwi_opts *opt, double *misfit, float *grad, float *illum_out,
                     float *resid_out) {
    return guarded(h, [&](Engine &e) {
        e.fwi_gradient(shots, nshots, dobs, mute, vpmask, opt, misfit, grad, illum_out, resid_out);
    });
}

int sfd_afwi_gradient(sfd_handle h, const sfd_shot *shots, int nshots, const float *dobs, const float *mute,
                      const sfd_afwi_opts *opt, double *misfit, float *grad, float *illum) {
    return guarded(h, [&](Engine &e) { e.afwi_gradient(shots, nshots, dobs, mute, opt, misfit, grad, illum); });
}

int sfd_bandpass(sfd_handle h, const sfd_bandpass_opts *opt, int nt, int ntr, float *traces) {
    if (!opt) return SFD_EINVAL;
    return guarded(h, [&](Engine &e) { e.bandpass(*opt, nt, ntr, traces); });
}

int sfd_mute_mask(sfd_handle h, const sfd_shot *shots, int nshots, float ot, const sfd_mute_opts *opt,
                  float *mask_out) {
    if (!opt) return SFD_EINVAL;
    return guarded(h, [&](Engine &e) { e.mute_mask(shots, nshots, ot, *opt, mask_out); });
}

// FWI inversion: Inversion.f90 l_bfgs with compute_fct_gdt as fctgdt
// (progs/afwi: the model is set, then misfit and gradient are computed).
int sfd_fwi_lbfgs(sfd_handle h, const sfd_shot *shots, int nshots, const float *dobs, const float *mute,
                  const float *vpmask, const sfd_fwi_opts *fo, const sfd_lbfgs_opts *lo,
                  const sfd_precond_opts *pc, float *x, sfd_lbfgs_report *rep) {
    if (!lo || !fo || !x) return SFD_EINVAL;
    return guarded(h, [&](Engine &e) {
        const long n = (long)e.g.nz * e.g.nx;
        std::vector<double> vpinit(x, x + n);
        std::unique_ptr<LbfgsHooks> hooks;
        if (pc && pc->enable) hooks.reset(make_cnn_precond(e.g.nz, e.g.nx, *pc));
        std::vector<float> il(hooks && hooks->wants_illumination() ? n : 0);
        LbfgsFct f = [&](const float *xx, float *g, double *fv) {
            e.set_model(xx, nullptr);
            e.raw_illum_host = il.empty() ? nullptr : il.data();
            e.fwi_gradient(shots, nshots, dobs, mute, vpmask, fo, fv, g, nullptr, nullptr);
            e.raw_illum_host = nullptr;
            if (!il.empty()) hooks->set_illumination(il.data(), n);
            return 0;
        };
        lbfgs_drive(*lo, n, x, vpmask, vpinit.data(), f, rep, hooks.get());
    });
}

int sfd_born_adjoint_exact(sfd_handle h, const sfd_shot *shots, int nshots, const float *dbar, float *mbar) {
    return guarded(h, [&](Engine &e) { e.born_adjoint_exact(shots, nshots, dbar, mbar); });
}
int sfd_adjoint_exact(sfd_handle h, const sfd_shot *shots, int nshots, const float *dbar, float *sbar,
                      int extract_kind) {
    return guarded(h, [&](Engine &e) { e.adjoint_exact(shots, nshots, dbar, sbar, extract_kind); });
}

int sfd_last_timing(sfd_handle h, double *ms, double *pts) {
    return guarded(h, [&](Engine &e) {
        if (ms)
            for (int k = 0; k < 4; k++) ms[k] = e.timing[k];
        if (pts) pts[0] = e.pts;
    });
}

int sfd_set_tile(sfd_handle h, int tile_z, int tile_x) {
    return guarded(h, [&](Engine &e) {
        if (tile_z <= 0) {
            e.autotune_pending = true;
            return;
        }
        if (!Engine::tile_supported(tile_z, tile_x)) throw EngineError(SFD_EUNSUPPORTED, "tile shape not compiled");
        e.tz = tile_z;
        e.tx = tile_x;
        // a strip width set for the previous shape that this one does not
        // compile falls back to the shape's default
        if (e.strip_px && !pick_kernel(e.tz, e.tx, false, false, 0, e.strip_px)) e.strip_px = 0;
        e.autotune_pending = false;
    });
}

int sfd_get_tile(sfd_handle h, int *tile_z, int *tile_x, int *strip_px, int *time_block, char *log, int len) {
    return guarded(h, [&](Engine &e) {
        if (time_block) *time_block = e.time_block;
        if (tile_z) *tile_z = e.tz;
        if (tile_x) *tile_x = e.tx;
        if (strip_px) *strip_px = e.strip_px ? e.strip_px : (e.tx == 32 ? 4 : 16);
        if (log && len > 0) snprintf(log, (size_t)len, "%s", e.tune_log.c_str());
    });
}

int sfd_copy_calibration(size_t nbytes, int reps, double *avg_ms) {
    if (nbytes < 16 || reps < 1 || !avg_ms) return SFD_EINVAL;
    const size_t n4 = nbytes / 16;
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = SFD_OK;
    if (hipMalloc(&a, n4 * 16) != hipSuccess || hipMalloc(&b, n4 * 16) != hipSuccess) rc = SFD_ENOMEM;
    if (rc == SFD_OK && (hipMemset(a, 0, n4 * 16) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
                         hipEventCreate(&e1) != hipSuccess))
        rc = SFD_EHIP;
    if (rc == SFD_OK) {
        const int grid = 4096;
        hipLaunchKernelGGL(copy_calib_kernel, dim3(grid), dim3(256), 0, 0, (const float4 *)a, (float4 *)b, n4);
        (void)hipEventRecord(e0, 0);
        for (int r = 0; r < reps; r++)
            hipLaunchKernelGGL(copy_calib_kernel, dim3(grid), dim3(256), 0, 0, (const float4 *)a, (float4 *)b, n4);
        (void)hipEventRecord(e1, 0);
        float ms = 0;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = SFD_EHIP;
        *avg_ms = ms / reps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

int sfd_stream_calibration(size_t npoints, int reps, double *avg_ms) {
    if (npoints < 1 || reps < 1 || !avg_ms) return SFD_EINVAL;
    float *p[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = SFD_OK;
    for (int k = 0; k < 6 && rc == SFD_OK; k++)
        if (hipMalloc(&p[k], npoints * 4) != hipSuccess || hipMemset(p[k], 0, npoints * 4) != hipSuccess)
            rc = SFD_ENOMEM;
    if (rc == SFD_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = SFD_EHIP;
    if (rc == SFD_OK) {
        const dim3 grid((unsigned)((npoints + 255) / 256));
        for (int r = 0; r <= reps; r++) {   // the first launch warms up, untimed
            if (r == 1) (void)hipEventRecord(e0, 0);
            hipLaunchKernelGGL(stream5_calib_kernel, grid, dim3(256), 0, 0, (const float *)p[0], (const float *)p[1],
                               (const float *)p[2], (const float *)p[3], (const float *)p[4], p[5], npoints);
        }
        (void)hipEventRecord(e1, 0);
        float ms = 0;
        if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            rc = SFD_EHIP;
        *avg_ms = ms / reps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (int k = 0; k < 6; k++)
        if (p[k]) (void)hipFree(p[k]);
    return rc;
}

int sfd_set_density(sfd_handle h, const float *rho) {
    return guarded(h, [&](Engine &e) { e.set_density(rho); });
}

int sfd_set_elevation(sfd_handle h, const float *elev) {
    return guarded(h, [&](Engine &e) { e.set_elevation(elev); });
}

int sfd_set_option(sfd_handle h, int opt, int value) {
    return guarded(h, [&](Engine &e) {
        switch (opt) {
        case SFD_OPT_GRAPHS: e.use_graphs = value != 0; break;
        case SFD_OPT_PLAN_REUSE: e.plan_reuse = value != 0; break;
        case SFD_OPT_GRAPH_CHUNK:
            if (value < 1) throw EngineError(SFD_EINVAL, "graph chunk must be >= 1");
            e.graph_chunk = value;
            break;
        case SFD_OPT_XCD_REMAP: e.xcd_remap = value != 0; break;
        case SFD_OPT_TIME_BLOCK:
            if (value != 1 && value != 2) throw EngineError(SFD_EINVAL, "time block must be 1 or 2");
            e.time_block = value;
            break;
        case SFD_OPT_STRIP_PX:
            // compiled pairs: 64x32 tiles with 4- or 8-column strips, 64x64 with 16
            if (value != 0 && !pick_kernel(e.tz, e.tx, false, false, 0, value))
                throw EngineError(SFD_EINVAL, "strip width " + std::to_string(value) + " is not compiled for " +
                                                  std::to_string(e.tz) + "x" + std::to_string(e.tx) +
                                                  " tiles (64x32: 4 or 8; 64x64: 16)");
            e.strip_px = value;
            break;
        case SFD_OPT_BATCH_MIB:
            if (value < 0) throw EngineError(SFD_EINVAL, "batch memory cap must be >= 0 MiB");
            e.batch_bytes_cap = (size_t)value << 20;
            break;
        case SFD_OPT_KERNEL_EVENTS:
            if (value < 0) throw EngineError(SFD_EINVAL, "event sampling must be >= 0");
            e.event_every = value;
            break;
        default: throw EngineError(SFD_EINVAL, "unknown option");
        }
    });
}

int sfd_last_waves(sfd_handle h, int *waves) {
    return guarded(h, [&](Engine &e) {
        if (waves) *waves = e.waves_last;
    });
}

int sfd_kernel_totals(sfd_handle h, double *us, double *bytes, double *pts, int *steps) {
    return guarded(h, [&](Engine &e) {
        for (int k = 0; k < 2; k++) {
            if (us) us[k] = e.kstat_us[k];
            if (bytes) bytes[k] = e.kstat_bytes[k];
            if (pts) pts[k] = e.kstat_pts[k];
            if (steps) steps[k] = e.kstat_n[k];
        }
    });
}

int sfd_kernel_stats(sfd_handle h, double *avg_us, int *nsamples) {
    return guarded(h, [&](Engine &e) {
        for (int k = 0; k < 2; k++) {
            if (avg_us) avg_us[k] = e.kstat_n[k] ? e.kstat_us[k] / e.kstat_n[k] : 0.0;
            if (nsamples) nsamples[k] = e.kstat_n[k];
        }
    });
}


}  // extern "C"
