// engine.h — host-side state of one seisfd handle.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/seisfd.h"
#include "common.h"
#include "fd_kernels.h"

// Grow-only device buffer reused across calls (no allocation in steady state).
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    void reserve(size_t n);
    ~DevBuf();
};

// Pinned host memory that grows on demand (device-to-host copies of images
// at full PCIe rate, no per-call allocation).
struct HostBuf {
    void *p = nullptr;
    size_t cap = 0;
    void reserve(size_t n);
    ~HostBuf();
};

// Host-side plan of one shot: window, padded-grid geometry, tiles, and the
// injection / extraction tables of the step kernel.
struct ShotPlan {
    float oxw = 0;
    int nxw = 0, nt = 0, ntr = 0, ntsnap = 0;
    int NZ = 0, NX = 0, LDZ = 0, zoff = 0, zsh = 0, xsh = 0, ntz = 0, ntx = 0, tile0 = 0, i0 = 0;
    int ntop = 0, iz0 = 0;         // FD rows above model row 1 (1 - nmin1) and the padded row of row 1
    std::vector<int> fs_row;       // free surface: per padded column, the padded row of ielev_z
    std::vector<float> fs_w;       // per padded column: sinc weights of delev_z / dz and an off-grid flag
    int snap_mask = 0;             // some window column's surface lies below model row 1
    int n_inj_traces = 1;
    std::vector<int> colsrc, colmap_raw;
    std::vector<float> taper;
    std::vector<int> inj_tile, inj_cell, inj_cptr, inj_tptr, inj_ttr;
    std::vector<float> inj_coef;   // 4 per contribution: v2r, sinc_z, sinc_x, local trace (int bits)
    std::vector<int> injb_tile, injb_cell, injb_cptr, injb_tptr, injb_ttr;
    std::vector<float> injb_coef;
    std::vector<int> inj_desc, inj_map, injb_desc, injb_map;   // step-kernel form of the two sets
    std::vector<int> ext_tile, ext_rec, ext_pos;
    std::vector<float> ext_w;
    int ext_kind = 0;
    void take_backward(const ShotPlan &o);
};

struct InjSet {
    const int *tile = nullptr, *cell = nullptr, *cptr = nullptr, *tptr = nullptr, *ttr = nullptr;
    const sfd::InjCoef *coef = nullptr;
    // region plan of the two-step kernel
    const int *tile2 = nullptr, *cell2 = nullptr, *cptr2 = nullptr, *tptr2 = nullptr, *ttr2 = nullptr;
    const sfd::InjCoef *coef2 = nullptr;
    int lo = 0, hi = 0;   // range of injecting tiles
    const int *desc = nullptr, *map = nullptr;
    void apply(sfd::ShotDev &S) const;
};

// Injection plan over the two-step kernel's regions (tile + 4-cell ring).
struct RegionPlan {
    std::vector<int> tile, cell, cptr, tptr, ttr;
    std::vector<float> coef;
    bool ok = true;   // false: a region needs more than the kernel's trace table
};

struct BatchNeeds {
    bool backward_injection = false;
    bool snapshots = false;
    bool snap_fp16 = false;
    bool snap_global = false;   // FP16 with one scale per snapshot (SFD_SNAP_FP16G)
    bool imaging = false;
    bool born = false;
    bool rhovp = false;         // a second image (the rho gradient of the v-rho imaging condition)
};

struct Batch {
    DevBuf mem;
    char *ext_base = nullptr;   // when set: the batch lives here (ext_cap bytes), mem is not used
    size_t ext_cap = 0;
    std::vector<RegionPlan> rp, rpb;   // two-step kernel region plans (forward / backward injections)
    int ns = 0, nblocks = 0;
    std::vector<sfd::ShotDev> shots_h;
    std::vector<InjSet> injb;
    std::vector<const int *> colsrc_d;
    std::vector<int> blk_shot_h;
    sfd::ShotDev *shots = nullptr;
    const int *blk_shot = nullptr;
    // Tiles are laid out shot by shot in order of decreasing record length
    // (order[k] = the shot in place k, nt_sorted[k] its nt, tiles_upto[k] the
    // tiles of places 0..k-1), so the shots still running at any step form a
    // prefix of the tiles: a launch covers only those (ragged batches).
    std::vector<int> order, nt_sorted, tiles_upto;
    const int *order_d = nullptr;
    float *snap_part = nullptr;   // per-shot partial maxima of the per-snapshot FP16 scale
    int active_shots(int n) const;   // shots with nt > n (step n is in their record)
    int active_tiles(int n) const { return tiles_upto[active_shots(n)]; }
    // Reuse by a later call with the same plan (empty: none): the key of the
    // plan the device tables hold, the shot descriptors as built and the
    // byte ranges zeroed at build time (everything but the snapshots).
    std::string key;
    std::vector<sfd::ShotDev> shots_built;
    std::vector<std::pair<size_t, size_t>> zero_ranges;
};

void second_difference(std::vector<float> &w, float dt);
void sfd_host_mksinc(float *sinc, int lsinc, float shift);

struct Engine {
    sfd_grid g;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    std::vector<float> vel_h, velb_h;
    std::vector<float> elev_h;            // free-surface elevation per model column (empty: 0 everywhere)
    std::vector<float> rho_h;             // density per model cell (empty: constant density)
    float *d_rho = nullptr;
    std::vector<std::vector<float>> src_shifted;   // variable density: half-step shifted source wavelets
    uint64_t model_version = 0;           // bumped by every set_model (plans read vel_h / velb_h)
    std::vector<ShotPlan> rtm_plans;      // host plan behind fwd_batch when its key is an RTM's
    float *d_vel = nullptr, *d_velb = nullptr;
    bool have_model = false;
    int tz = 64, tx = 32;
    bool autotune_pending = false;
    std::string tune_log;                 // candidates and times of the last autotune
    int xcd_remap = 1;
    int strip_px = 0;        // columns per wavefront (0: tile width / 4)
    int time_block = 1;      // steps per launch: 1, or 2 (two-step kernel where it applies)
    bool plan_reuse = true;  // SFD_OPT_PLAN_REUSE
    int graph_chunk = 200;   // steps per replayed block (rounded up to a multiple of lcm(kLevels, snapi))
    bool use_graphs = true;
    int event_every = 1;
    std::vector<hipEvent_t> kev;          // start/stop of every block of steps
    DevBuf nbase;                         // device-held first step of the running block
    // per pass (0 forward time, 1 reverse time): event-timed microseconds,
    // steps, FD point-updates and algorithmic HBM bytes of the timed launches
    double kstat_us[2] = {0, 0};
    int kstat_n[2] = {0, 0};
    double kstat_pts[2] = {0, 0}, kstat_bytes[2] = {0, 0};
    double timing[4] = {0, 0, 0, 0};
    double pts = 0;
    float *raw_illum_host = nullptr;      // fwi_gradient: also copy the summed illumination here
    bool tb_ok = true;                    // the current batch's region plans fit the two-step kernel
    Batch fwd_batch;
    // allocation-free forward modeling: its batch in the library's device
    // scratch array (this handle's device's copy), eager untimed launches
    Batch noalloc_batch;
    char *scratch_dev = nullptr;
    std::string scratch_err;
    bool eager_untimed = false;
    DevBuf tmp_mem, fwi_mem, acc_mem;
    size_t batch_bytes_cap = 0;           // SFD_OPT_BATCH_BYTES (0: free device memory decides)
    int waves_last = 0;                   // shot waves of the last call
    HostBuf img_host;
    DevBuf resident;                      // traces kept in HBM by sfd_set_traces
    size_t resident_n = 0;
    void set_traces(const float *tr, size_t n);

    explicit Engine(const sfd_grid &grid);
    ~Engine();
    void set_model(const float *vel, const float *vel_bwd);
    void forward(const sfd_shot *shots, int ns, float *traces_out, int extract_kind, Batch *ext = nullptr);
    void rtm(const sfd_shot *shots, int ns, const float *traces_in, float *image, float *illum, int snap_mode,
             bool lsrtm);
    void born(const sfd_shot *shots, int ns, const float *refl, float *traces_out);
    void adjoint_exact(const sfd_shot *shots, int ns, const float *dbar, float *sbar, int extract_kind);
    void born_adjoint_exact(const sfd_shot *shots, int ns, const float *dbar, float *mbar);
    void mute_mask(const sfd_shot *shots, int ns, float ot, const sfd_mute_opts &o, float *mask_out);
    void bandpass(const sfd_bandpass_opts &o, int nt, int ntr, float *traces);
    void fwi_gradient(const sfd_shot *shots, int ns, const float *dobs, const float *mute, const float *vpmask,
                      const sfd_fwi_opts *opt, double *misfit, float *grad, float *illum_out, float *resid_out);
    void afwi_gradient(const sfd_shot *shots, int ns, const float *dobs, const float *mute,
                       const sfd_afwi_opts *opt, double *misfit, float *grad, float *illum_out);

    static bool tile_supported(int TZ, int TX);
    void check_shots(const sfd_shot *shots, int nshots, int snaps = -1, bool may_tune = true);
    void autotune(const sfd_shot *shots, int ns, int snaps);
    float padded_vel(const std::vector<float> &vh, const ShotPlan &p, int icz, int icx) const;
    void plan_shot(const sfd_shot &s, ShotPlan &p) const;
    void plan_injection(ShotPlan &p, const std::vector<float> &vh, int nsrc, const float *sz, const float *sx,
                        int kind) const;
    bool surface() const { return g.surf_type != 0; }
    bool has_rho() const { return !rho_h.empty(); }
    void set_density(const float *rho);
    const float *source_wavelet(const sfd_shot &sh, int k);
    float padded_rho(const ShotPlan &p, int icz, int icx) const;
    bool ghosts_or_surface() const { return g.surf_type != 0 || g.shot_type != 0 || g.rec_type != 0; }
    void set_elevation(const float *elev);
    void plan_extraction(ShotPlan &p, const std::vector<float> &vh, const sfd_shot &s, int kind) const;
    void build_batch(Batch &B, std::vector<ShotPlan> &plans, const BatchNeeds &need);
    void reset_batch(Batch &B);
    std::string plan_key(const char *kind, const sfd_shot *shots, int ns, int a, int b) const;
    void upload_shots(Batch &B);
    void prepare_models(Batch &B, std::vector<ShotPlan> &plans, const float *d_model);
    void zero_wavefields(Batch &B, std::vector<ShotPlan> &plans);
    sfd::StepArgs base_args(const Batch &B) const;
    void launch_step(const Batch &B, sfd::StepArgs A, bool adj);
    void run_loop(const Batch &B, sfd::StepArgs A, bool adj, int nsteps);
    void launch_step2(const Batch &B, sfd::StepArgs A, bool adj);
    double step_bytes(const Batch &B, const sfd::StepArgs &A, int n, bool two) const;
    bool use_step2(const sfd::StepArgs &A) const;
    void reset_kstats();
    double count_pts(const std::vector<ShotPlan> &plans) const;
    void upload_inj_traces(Batch &B, int s, const float *tr, size_t count);
    void traces_time_major(const Batch &B);
    void accumulate_images(Batch &B, std::vector<ShotPlan> &plans, float *d_img, float *d_ill,
                           float *d_img2 = nullptr);
    size_t shot_bytes(const sfd_shot &s, const BatchNeeds &need) const;
    std::vector<std::pair<int, int>> plan_waves(const sfd_shot *shots, int ns, const BatchNeeds &need);
};
