/*
 * seisfd.h — C-ABI of the MI355X acoustic finite-difference engine.
 *
 * Every entry point is plain C: integers, floats and host pointers, no C++
 * or torch types, so Fortran (iso_c_binding), C and Python (ctypes) callers
 * bind to it directly.  All arrays are column-major with depth (z) fastest,
 * i.e. the layout of the reference's `vel(nz,nx)` and `trace(nt,ntr)`.
 *
 * Return codes: 0 on success, a negative SFD_E* code on failure (the message
 * is available from sfd_last_error()).  Nothing aborts the process; the
 * reference calls `erexit` instead (libs/libwave/ReadSouVelRho.f90:54,385,675).
 *
 * Threading: one handle per host thread.  A handle owns its device buffers,
 * its HIP stream and its captured graphs; the caller owns every host buffer.
 */
#ifndef SEISFD_H
#define SEISFD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SFD_OK            0
#define SFD_EINVAL       -1   /* bad argument / geometry                      */
#define SFD_ENOMEM       -2   /* device or host allocation failed             */
#define SFD_EHIP         -3   /* HIP runtime error                             */
#define SFD_EUNSUPPORTED -4   /* option outside the implemented subset        */
#define SFD_ESTATE       -5   /* call order (e.g. no model set)               */
#define SFD_EUNSTABLE    -6   /* Courant check failed (ReadSouVelRho.f90:629-675) */
#define SFD_ENOTFINITE   -7   /* NaN/Inf met in a misfit or wavefield guard   */
#define SFD_EIO          -8   /* state file missing, unreadable or not matching */
#define SFD_EALIASED     -9   /* vel_check sampling stop: vmin / fmax < 2 max(dz,dx)
                                 (ReadSouVelRho.f90:688-752)                  */

/* Snapshot storage for RTM / FWI source wavefields. */
#define SFD_SNAP_FP32     0   /* reference behaviour: FP32 copies           */
#define SFD_SNAP_FP16S    1   /* FP16, one max|S| scale per snapshot and 64 x 4 cell block */
#define SFD_SNAP_FP16G    2   /* FP16, one max|S| scale per snapshot (the whole window) */

/* Receiver extraction semantics (libs/libwave/Extraction.f90). */
#define SFD_EXTRACT_SINC_V2    0  /* Extraction_array_sinc: gain v^2 (:109-134)   */
#define SFD_EXTRACT_NEAREST    1  /* Extraction_array_simple_afwi_noomp (:71-106) */
#define SFD_EXTRACT_SINC_LSRTM 2  /* Extraction_array_LSRTM_sinc, gain 1/(dz dx) (:217-246) */
#define SFD_EXTRACT_SINC_UNIT  3  /* Extraction_array_sinc_afwi: gain 1 (:137-162) */

/*
 * Grid / run parameters.  Mirrors the fields of the reference's
 * GeneralParam (libs/libwave/GeneralParamTypes.f90:9-48) that the 2D
 * constant-density path reads, plus the model header (n1,n2,o1,o2,d1,d2).
 */
typedef struct sfd_grid {
    int32_t nz, nx;          /* model samples: depth (fast axis), lateral     */
    float   oz, ox;          /* model origin (m)                              */
    float   dz, dx;          /* spacing (m)                                   */
    float   dt;              /* time step (s)                                 */
    int32_t ntaper;          /* absorbing pad per side (cells), >= 10         */
    int32_t lsinc;           /* sinc interpolator length (7)                  */
    int32_t snapi;           /* snapshot interval (steps)                     */
    float   aperture_x;      /* migration aperture (m), ExtractPadModel.f90:55 */
    int32_t rec_type;        /* != 0: receivers read u(r) - u(mirror of r)
                                (Extraction.f90:46-48,88-90,548-556); the
                                simple receiver injection of FWI's adjoint
                                pass adds the mirror source (Injection.f90:583) */
    int32_t shot_type;       /* != 0: sinc injections (sources, RTM/LSRTM
                                receivers) add the mirror source
                                (Injection.f90:432,447-450); > 0 also extends
                                the FD rows above the model to ntaper+5
                                (ExtractPadModel.f90:229)                     */
    int32_t surf_type;       /* != 0: free surface at the elevation of
                                sfd_set_elevation, set on u2 every step
                                (Boundary.f90:1609-1658), snapshots zero above
                                it (Extraction.f90:580-592); > 0 also drops
                                the top Higdon band and keeps lsinc FD rows
                                above the model (Boundary.f90:104-124,
                                ExtractPadModel.f90:230).  The exact adjoints
                                (sfd_adjoint_exact, sfd_born_adjoint_exact)
                                need all three types 0.                       */
    float   fmax;            /* source's highest frequency (Hz) for vel_check's
                                sampling stop; 0: the reference default 30
                                (ReadParam.f90:18)                            */
} sfd_grid;

/*
 * One shot: a source, its wavelet and its receiver spread.  Replaces the
 * reference's sourcevec/datavec TraceSpace arrays
 * (libs/libwave/DataSpaceTypes.f90:110-117): coord(1)=z, coord(2)=x.
 */
typedef struct sfd_shot {
    float        sz, sx;     /* source depth / lateral position (m)           */
    int32_t      nt;         /* record length (samples at dt)                 */
    int32_t      ntr;        /* receivers                                     */
    const float *wavelet;    /* nt samples                                    */
    const float *rz;         /* ntr receiver depths (m)                       */
    const float *rx;         /* ntr receiver lateral positions (m)            */
} sfd_shot;

/* Data misfit norms (fct_compute / gdt_compute, libs/libinv/Norm.f90:12-141),
 * numbered as the reference's dat_nrm_type (Inversion_types.f90:61-67). */
#define SFD_NORM_L2         2   /* f = sum(r^2)/2 in FP64; adjoint source r        */
#define SFD_NORM_L1         1   /* f = sum|r|; adjoint source sign(r)              */
#define SFD_NORM_HUBER     12   /* |r|-t/2 or r^2/(2t); clamp(r/t, -1, 1)          */
#define SFD_NORM_CAUCHY     3   /* log(1+(r/t)^2)/2; r/(t^2+r^2)                   */
#define SFD_NORM_HYPERBOLIC 4   /* sqrt(t^2+r^2)-t; (r/t)/sqrt(1+r^2/t^2) (Norm.f90
                                   only: the reference's parameter parsing cannot
                                   select it)                                      */

/* FWI gradient post-processing options (compute_function_gradient.f90:332-343). */
typedef struct sfd_fwi_opts {
    float   illupow;         /* illumination power, Inversion_types.f90:60    */
    int32_t rect_z, rect_x;  /* triangle smoothing half-lengths (cells)       */
    int32_t do_z, do_x;      /* smoothing switches (Smoothing.f90:57-61)      */
    int32_t snap_mode;       /* SFD_SNAP_*                                    */
    int32_t norm;            /* SFD_NORM_* (0 = L2), data_nrm_type             */
    float   norm_thresh;     /* data_threshold t (Inversion_types.f90:62); 0 =
                                per shot, max|r| of the first trace whose
                                residual is not all zero, as Compute_OF_RES_ADJ
                                passes its one THRESH to every trace
                                (Objective_Function_Residual_Adjsource.f90:26-38;
                                Norm.f90:36-40).  The non-L2 norms sum each
                                trace's FP32 terms in FP32, in time order.    */
} sfd_fwi_opts;

typedef struct sfd_ctx *sfd_handle;

/* ---- lifetime (replaces propagator_acoustic's per-call allocation,
 *      libs/libwave/Propagator.f90:144-160,261-262) ------------------------ */
int  sfd_create(sfd_handle *h, const sfd_grid *g);
int  sfd_destroy(sfd_handle h);
const char *sfd_last_error(sfd_handle h);
int  sfd_device_info(char *name, int len, int *cu_count, size_t *hbm_bytes);

/*
 * Model: unpadded vel(nz,nx) in m/s, column-major.  Windowing and
 * edge-replication padding are done on the device per shot, as
 * read_window_vel / copy_window_vel_gath + model_pad do
 * (libs/libwave/ExtractPadModel.f90:87-285,395-507;
 *  libs/libwave/ReadSouVelRho.f90:545-596).  Courant check as vel_check
 * (ReadSouVelRho.f90:603-675).  vel_bwd may be NULL (the reference's optional
 * `vel2` used by the RTM backward pass, to_disk_to_memory_RTM.f90:438).
 */
int  sfd_set_model(sfd_handle h, const float *vel, const float *vel_bwd);

/*
 * Variable density (genpar%withRho, the model's rho file read by
 * read_window_vel, libs/libwave/ExtractPadModel.f90:259-270): rho(nz,nx) > 0,
 * column-major; NULL (the default) is constant density.  With a density the
 * engine propagates with FD_2D_derivatives_acoustic_forward_grid
 * (FDderivatives.f90:588-657: staggered gradients times 1/rho2, rho2 the
 * density sinc-interpolated to the half cells by Interpolate, Interpolate.f90:
 * 44-130, divergence times v^2 rho, coefficients FD_acoustic_rho_init_coefs
 * over delta), injects with gain v^2 rho (Injection.f90:172-176) and shifts
 * each source wavelet half a time step forward as readsou does
 * (ReadSouVelRho.f90:328-349), in sfd_forward and sfd_rtm (both RTM passes use
 * that operator, to_disk_to_memory_RTM.f90:389-397,454-462).  sfd_lsrtm,
 * sfd_born, sfd_fwi_gradient and the exact adjoints return SFD_EUNSUPPORTED
 * while a density is set; tiles must be 64 x 32 with 4-column strips.
 */
int  sfd_set_density(sfd_handle h, const float *rho);

/*
 * Free-surface elevation (m, depth axis) per model column, nx values; NULL
 * (the default) is 0 everywhere, which is what the reference's drivers set
 * (progs/rtm/Acoustic_RTM_sep.f90:118-119, 3DRTM.f90:74-75).  Replaces
 * ModelSpace_elevation%elev and ModelSpace_elevation_parameters
 * (libs/libwave/ModelSpaceTypes.f90:68-73,128-141): the surface cell
 * ielev_z = nint((elev - oz)/dz) + 1 must lie in 1..nz.  FD columns outside a
 * shot's window take the nearest window column's elevation.  Read only when
 * surf_type != 0.
 */
int  sfd_set_elevation(sfd_handle h, const float *elev);

/*
 * Forward modeling of nshots shots (Acoustic_modeling_sep.f90:80-88 for one
 * shot; to_disk_to_memory_AMOD.f90:57-65 with extract=NEAREST).  Traces are
 * written per shot, each shot's block (nt, ntr) time-fastest, blocks
 * concatenated in shot order.  Source injection is Injection_sinc
 * (Injection.f90:151-194,402-456).
 */
int  sfd_forward(sfd_handle h, const sfd_shot *shots, int nshots,
                 float *traces_out, int extract_kind);

/*
 * Allocation-free forward modeling (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * results and errors as sfd_forward, but the call allocates nothing: the
 * batch (wavefield levels, model copies, plan tables, traces) lives in the
 * library's device scratch array of 64 MiB (one per GPU, held until the
 * process ends); launches are eager and untimed (no graphs, events or
 * autotuning).  A batch larger than the array is SFD_ENOMEM with the sizes in
 * the message.  The array is shared by every call on a GPU: calls that
 * overlap on one GPU race.  The model must be set (sfd_set_model) before.
 */
int  sfd_forward_noalloc(sfd_handle h, const sfd_shot *shots, int nshots,
                         float *traces_out, int extract_kind);

/*
 * RTM of one or more shots (to_disk_to_memory_RTM.f90:358-492): forward with
 * snapshots every snapi steps, reference-style backward with the adjoint
 * stencil (FDderivatives.f90:847-891) injecting traces_in with
 * Injection_sinc, zero-lag imaging and illumination
 * (Imaging.f90:228-275).  image/illum are full-model (nz,nx) and accumulated
 * with += through the nearest-index scatter of mod_copy_image
 * (ExtractPadModel.f90:662-695).
 */
int  sfd_rtm(sfd_handle h, const sfd_shot *shots, int nshots,
             const float *traces_in, float *image, float *illum, int snap_mode);

/*
 * Keep traces resident in HBM (no reference counterpart): count floats, the
 * shots' (nt, ntr) traces concatenated in shot order as sfd_rtm takes them.
 * sfd_rtm / sfd_lsrtm called with traces_in = NULL (sfd_fwi_gradient with
 * dobs = NULL) then read this copy
 * instead of transferring host data each call (repeated imaging, LSRTM
 * iterations); the count must equal the call's sum of nt * ntr.  traces =
 * NULL or count = 0 releases the copy.
 */
int  sfd_set_traces(sfd_handle h, const float *traces, size_t count);

/*
 * Born modeling (to_disk_to_memory_BORN.f90:123-220): pass 1 forward with
 * snapshots (source second-differenced as readsou does with Born=.true.,
 * ReadSouVelRho.f90:318-326), pass 2 scattering from refl(nz,nx) via
 * Injection_Born (Imaging.f90:747-788), extraction with gain 1/(dz dx).
 */
int  sfd_born(sfd_handle h, const sfd_shot *shots, int nshots,
              const float *refl, float *traces_out);

/*
 * Exact discrete adjoint of sfd_forward (no reference counterpart: the
 * reference's reverse-time pass, as in sfd_rtm, runs the adjoint-style
 * stencil forward-in-reverse with the absorbing boundary re-applied, which is
 * not the transpose of the forward operator).  With F the linear map from a
 * shot's wavelet to its traces (sinc injection with gain v^2, 5-level update,
 * Higdon bands, extraction of extract_kind), sbar = F^T dbar: dbar (nt, ntr)
 * per shot concatenated as sfd_forward writes traces, sbar nt per shot.
 * <F s, d> = <s, F^T d> holds to FP32 rounding (the dot-product test).
 */
int  sfd_adjoint_exact(sfd_handle h, const sfd_shot *shots, int nshots,
                       const float *dbar, float *sbar, int extract_kind);

/*
 * Exact discrete adjoint of sfd_born (no reference counterpart: the
 * reference's LSRTM imaging, sfd_lsrtm, is the reverse-time approximation of
 * it).  With B the linear map from refl(nz,nx) to the traces of sfd_born
 * (background pass fixed), mbar = B^T dbar summed over the shots: dbar (nt,
 * ntr) per shot concatenated as sfd_born writes traces, mbar (nz, nx)
 * overwritten.  <B m, d> = <m, B^T d> holds to FP32 rounding.
 */
int  sfd_born_adjoint_exact(sfd_handle h, const sfd_shot *shots, int nshots,
                            const float *dbar, float *mbar);

/* LSRTM imaging (to_disk_to_memory_RTM.f90:494-579, Imaging.f90:698-745). */
int  sfd_lsrtm(sfd_handle h, const sfd_shot *shots, int nshots,
               const float *traces_in, float *image, float *illum);

/* Butterworth band-pass (BandPassParam, Bandpass.f90:13-47; the reference's
 * defaults: fhi 20 Hz, flo 0, nphi 14, nplo 14, phase 0). */
typedef struct sfd_bandpass_opts {
    float   flo, fhi;        /* cut-offs (Hz); flo dt <= 1e-6 skips the low-cut,
                                fhi dt >= 4.999999 the high-cut               */
    int32_t nplo, nphi;      /* poles (halved at zero phase), <= 16 sections  */
    int32_t phase;           /* 0 zero phase (forward and reversed), 1 minimum */
} sfd_bandpass_opts;

/*
 * bandpass (Bandpass.f90:89-147, the filters of libs/libwave/bandpass.c:72-266)
 * on ntr traces of nt samples at the grid's dt, in place, trace after trace
 * (time fastest).  BandpassSouTraces (:50-88) runs it on every shot's
 * wavelet and traces before an inversion.  Each trace is padded by 10
 * samples each side with sin^2 tapers of its end values, low-cut, high-cut,
 * and the nt samples after the front pad are kept.  (The reference's last
 * statement assigns data(npad+1:nt) to an nt-sample trace, a shape mismatch
 * that flang -O2 and compilers looping over the left-hand side read as
 * data(npad+1:npad+nt); flang -O0 stops there.)  Neither cut requested is
 * SFD_EINVAL (the reference's erexit).
 */
int  sfd_bandpass(sfd_handle h, const sfd_bandpass_opts *opt, int nt, int ntr, float *traces);

/* Offset-dependent top and bottom mutes (MuteParam, Mute.f90:14-46; the
 * reference's defaults: v0 1500 m/s, t0 0 s, v1 10000 m/s, t1 40 s, tramp
 * 0.5 s, Init_MuteParam :30-34). */
typedef struct sfd_mute_opts {
    float v0, t0;            /* top mute at sqrt(t0^2 + (h/v0)^2)              */
    float v1, t1;            /* bottom mute at sqrt(t1^2 + (h/v1)^2)           */
    float tramp;             /* taper length (s): sin^2 rise, cos^2 fall        */
} sfd_mute_opts;

/*
 * The mute mask of every trace (MuteParam_compute_mask, Mute.f90:50-98, with
 * compute_top_mute :171-195 and compute_bottom_mute :197-219): per shot an
 * (nt, ntr) block, blocks concatenated in shot order as sfd_fwi_gradient's
 * `mute` argument takes them.  The offset is h = x_receiver - x_source (the
 * reference adds the 3D y offset, 0 in 2D); ot is the traces' time origin
 * (s) and the sampling interval is the grid's dt.
 */
int  sfd_mute_mask(sfd_handle h, const sfd_shot *shots, int nshots, float ot, const sfd_mute_opts *opt,
                   float *mask_out);

/*
 * FWI misfit and gradient for a batch of shots, the body of compute_fct_gdt
 * (progs/afwi/compute_function_gradient.f90:208-350): per shot AMOD
 * (nearest extraction), residual w(dobs-dmod), misfit and adjoint source
 * under opt->norm (Objective_Function_Residual_Adjsource.f90:17-40,
 * libs/libinv/Norm.f90:12-141; L2: FP64, dprod),
 * AGRAD (simple injection, FWI imaging Imaging.f90:322-422), scatter-add,
 * then illumination normalisation, mask and triangle smoothing
 * (Smoothing.f90:87-227).  mute and vpmask may be NULL (all ones).
 * grad is the reference's `grad` (its sign convention: l_bfgs passes -grad
 * to LBFGS, Inversion.f90:128).  resid_out may be NULL.  dobs may be NULL:
 * the observed traces are then the copy kept in HBM by sfd_set_traces.
 */
int  sfd_fwi_gradient(sfd_handle h, const sfd_shot *shots, int nshots,
                      const float *dobs, const float *mute, const float *vpmask,
                      const sfd_fwi_opts *opt, double *misfit, float *grad,
                      float *illum_out, float *resid_out);

/* Options of sfd_afwi_gradient (the InversionParam fields 3DAFWI reads). */
typedef struct sfd_afwi_opts {
    int32_t snap_mode;       /* SFD_SNAP_*                                    */
    int32_t norm;            /* SFD_NORM_* (0 = L2), data_nrm_type             */
    float   norm_thresh;     /* data_threshold (0: per shot from the first trace
                                whose residual is not all zero, as in
                                sfd_fwi_gradient)                              */
    int32_t vprho_param;     /* with a density: 0 gradients for (vp, rho), 1 for
                                (vp, Ip = rho vp) (InversionParam%vprho_param,
                                ExtractPadModel.f90:643-646)                   */
} sfd_afwi_opts;

/*
 * The one-shot gradient of 3DAFWI.x (progs/afwi/3DAFWI.f90), AFWI_to_memory
 * (progs/afwi/to_disk_to_memory_AFWI.f90:229-415), for a batch of shots in
 * 2D: per shot forward modeling with Injection_sinc (readsou's half-step
 * shift with a density) and Extraction_array_sinc_afwi (gain 1) storing
 * snapshots; Compute_OF_RES_3D (Objective_Function_Residual_Adjsource.f90:
 * 42-67) under opt->norm; reverse time with the adjoint-style stencil
 * (the density operator with a density), Injection_sinc of the residual and
 * Imaging_condition_AFWI_from_memory (Imaging.f90:371-422), or with a density
 * Imaging_condition_AFWI_RHOVP_from_memory (:424-512); each window added into
 * the model (mod_copy_gradient_to_disk, ExtractPadModel.f90:592-660).
 * Outputs, summed over the shots (3DAFWI writes one shot's): misfit the sum
 * of the traces' norms (no scaling); grad (nz,nx) for vp, followed with a
 * density by (nz,nx) for rho (vprho_param 0) or Ip (1, the change of
 * variables on the model's vel and rho); illum (nz,nx), may be NULL.  dobs
 * may be NULL (the copy kept by sfd_set_traces); mute may be NULL.  With a
 * density every shot's window needs two columns (the imaging's one-sided x
 * difference).
 */
int  sfd_afwi_gradient(sfd_handle h, const sfd_shot *shots, int nshots, const float *dobs, const float *mute,
                       const sfd_afwi_opts *opt, double *misfit, float *grad, float *illum);

/*
 * Per-phase timing of the last call (the reference's verbose timers,
 * Propagator.f90:165-258): ms[0]=total GPU time of the propagation kernels,
 * ms[1]=forward pass, ms[2]=backward pass, ms[3]=host-side setup.
 * pts[0]= FD point-updates executed (sum over shots of nz_FD*nx_FD*nt).
 */
int  sfd_last_timing(sfd_handle h, double *ms, double *pts);

/*
 * Step kernel tile configuration (tile_z x tile_x cells per workgroup):
 * 64x32 (default) or 64x64.  tile_z <= 0 asks the autotuner to choose the
 * tile shape, wavefront strip width and time block (SFD_OPT_TIME_BLOCK) on
 * the next modeling call, by timing a short forward run of every compiled
 * combination on that call's batch.
 * sfd_get_tile reports the shape in use and, after an autotune, the timings.
 */
int  sfd_set_tile(sfd_handle h, int tile_z, int tile_x);
int  sfd_get_tile(sfd_handle h, int *tile_z, int *tile_x, int *strip_px, int *time_block, char *log,
                  int len);

/* Engine options (no reference counterpart: launch policy of the HIP path). */
#define SFD_OPT_GRAPHS       1  /* 1: replay blocks of time steps as one HIP graph (default) */
#define SFD_OPT_GRAPH_CHUNK  2  /* time steps per block (default 200, rounded up to a multiple of lcm(6, snapi)) */
#define SFD_OPT_XCD_REMAP    3  /* 1: XCD-contiguous tile order (default)          */
#define SFD_OPT_KERNEL_EVENTS 4 /* 1: time every block of steps with HIP events (default) */
#define SFD_OPT_STRIP_PX     5  /* columns per wavefront in the step kernel: 0 auto (4 for 64x32 tiles), 4, 8, 16 */
#define SFD_OPT_TIME_BLOCK   6  /* time steps per launch: 1 (default) or 2 (64x32 tiles; not Born passes) */
#define SFD_OPT_PLAN_REUSE   7  /* 1: an RTM/LSRTM call with the same shots, model and tile settings as
                                   the previous call reuses its plan tables on the device (default) */
#define SFD_OPT_BATCH_MIB    8  /* device memory (MiB) one wave of shots may use; 0 (default): the
                                   memory free on the device.  Calls whose shots do not fit run them
                                   in consecutive waves (shot order), with results equal to one batch */
int  sfd_set_option(sfd_handle h, int opt, int value);

/* Shot waves the last modeling call ran in (1: one batch).  The reference
 * runs shots one per OpenMP thread (compute_function_gradient.f90:246-256);
 * a wave is the batch of consecutive shots that fits in device memory. */
int  sfd_last_waves(sfd_handle h, int *waves);

/*
 * Step-kernel durations of the last call, from HIP events recorded on the
 * engine's stream around every block of back-to-back step launches (block
 * time / steps in the block): avg_us[0] forward-time pass, avg_us[1]
 * reverse-time pass; nsamples = steps timed (every step of the pass).
 */
int  sfd_kernel_stats(sfd_handle h, double *avg_us, int *nsamples);

/*
 * Totals of the same event-timed launches, per pass ([0] forward time, [1]
 * reverse time): us = kernel microseconds, bytes = algorithmic HBM bytes (24
 * per FD point-update, plus the snapshot write on forward snapshot steps and
 * the snapshot read(s) and 16 B of image and illumination read-modify-write
 * per window cell on reverse imaging steps), pts = FD point-updates, steps.
 * bytes / us is the achieved bandwidth of the pass's launches.  No reference
 * counterpart (the reference's per-phase timers, Propagator.f90:165-258).
 */
int  sfd_kernel_totals(sfd_handle h, double *us, double *bytes, double *pts, int *steps);

/*
 * Measurement helper (no reference counterpart): copies nbytes between two
 * device buffers with a plain streaming kernel (copy_calib_kernel) `reps`
 * times and returns the average kernel time.  Used to calibrate hardware
 * byte counters (FETCH_SIZE/WRITE_SIZE) against a known traffic volume.
 */
int  sfd_copy_calibration(size_t nbytes, int reps, double *avg_ms);

/*
 * Measurement helper (no reference counterpart): the streaming ceiling of a
 * step.  Runs a kernel that reads five FP32 arrays of npoints values and
 * writes a sixth (24 bytes per point, the step kernel's algorithmic traffic
 * and nothing else) `reps` times after one untimed launch; returns the
 * average kernel time.  Allocates and frees 24 * npoints bytes.
 */
int  sfd_stream_calibration(size_t npoints, int reps, double *avg_ms);

/* ---- L-BFGS (libs/libinv/lbfgs.f:49-558 LBFGS + MCSRCH + MCSTEP), reverse
 *      communication exactly as the reference's IFLAG protocol.  Vectors
 *      live on the device; the two-loop recursion and the line-search dot
 *      products are FP64 device reductions. ------------------------------- */
typedef struct sfd_lbfgs *sfd_lbfgs_handle;
int  sfd_lbfgs_create(sfd_lbfgs_handle *lb, int n, int m, double xmin, double xmax,
                      const float *mask, int boundt, int freeze_soft,
                      const double *xinit);
/* One call of LBFGS(N,M,X,F,G,...): x (in/out), f, g as the reference passes
 * them (g is the gradient of f); iflag/info as in lbfgs.f. */
int  sfd_lbfgs_step(sfd_lbfgs_handle lb, double *x, double f, const double *g,
                    int *iflag, int *info);
int  sfd_lbfgs_destroy(sfd_lbfgs_handle lb);

/* ---- Out-of-core L-BFGS (libs/libnlinv/nlinv_lbfgs.f: LBFGSOC :31-599,
 *      MCSRCHOC :614-967, FREEZE_XOC :1-22), driven one model update per call
 *      as progs/inversion/L_BFGS.f90:54-98 drives it: the complete optimiser
 *      state (the SAVE scalars of LBFGSOC and MCSRCHOC, the work array W and
 *      DIAG) lives in a file between calls, so every call may be a new
 *      process.  The vectors are FP64 on the device during the call. ------ */
typedef struct {
    int     m;           /* MSAVE (nlinv_read_array_params.f90:218 sets 5)       */
    int     nparam;      /* parameter blocks of n / nparam values (FREEZE_XOC)   */
    const double *xmin;  /* nparam lower bounds                                  */
    const double *xmax;  /* nparam upper bounds                                  */
    int     stp1_opt;    /* 1: first step |x|/|g|/100; otherwise 1/|g| (:336-341) */
    double  step_init;   /* first step when not 999999 (:342-344)                */
    int     clip_type;   /* BOUNDT: 1 clip every trial model (MCSRCHOC :850),
                            2 every accepted model (:521-524), 0 never           */
} sfd_lbfgsoc_opts;

/* One call of LBFGSOC(N,M,X,F,G,.FALSE.,DIAG,IPRINT,EPS,XTOL,STEPINIT,W,IFLAG,
 * MYINFO,...,XMIN,XMAX,NPARAM,SPT1_OPT,BOUNDT) with EPS = 1e-20 and XTOL =
 * machine epsilon (nlinv_read_array_params.f90:227-228).
 *   iflag (in/out): 0 starts a run (state_in is not read); 1 continues the run
 *     whose state state_in holds.  On exit: 1 = evaluate f, g at the new x;
 *     0 = converged; < 0 = LBFGSOC's error codes (-1 line search failed).
 *   info (in/out): MYINFO, set to 1 by a call that completed an iteration and
 *     otherwise left as passed (L_BFGS.f90 keeps it in Info.dat).
 *   x (n, in/out): the model f and g were evaluated at; on exit the next model.
 *   state_out receives the new state unless iflag < 0 on exit (written to a
 *   temporary file, then renamed); it may name the same file as state_in.
 * The state file also holds n, m and the bounds of the run: a mismatch with
 * the arguments is SFD_EIO, and so is a file whose closing 64-bit checksum
 * does not match its contents (truncated or damaged). */
int  sfd_lbfgsoc_step(const char *state_in, const char *state_out, const sfd_lbfgsoc_opts *o,
                      int n, double *x, double f, const double *g, int *iflag, int *info);

/* One run of progs/inversion/L_BFGS.f90 with lbfgs_type 1 (:54-98) on the
 * reference's own state files in directory dir, so that a job may alternate
 * between that program and this call, or switch mid-inversion, and resume
 * from the other's files.  Replaces the program's file handling
 * (nlinv_read_array_params.f90:139-248, L_BFGS.f90:69-90) and LBFGSOC's and
 * MCSRCHOC's state I/O (nlinv_lbfgs.f:275-281, :391-397, :509-565, :755-760,
 * :854-905); the SEP files of x, f and g stay with the caller, who passes
 * their values (the reference stores them as FP32).
 *   Reads:  Iflag.dat, Info.dat (list-directed integers; absent = 0); for an
 *           IFLAG other than 0: Diag.dat and WorkArray.dat (one sequential
 *           unformatted record each, DIAG = n doubles, W = n(2m+1)+2m doubles:
 *           q, rho, alpha, S, Y), LbfgsDat.dat (LBFGSOC's 33 scalars) and,
 *           when its INFO is not 0 and IFLAG is 1, McsrchDat.dat (MCSRCHOC's
 *           25 scalars).  IFLAG 2 (DIAGCO re-entry) is SFD_EINVAL.
 *   Writes: LbfgsDatOut.dat and McsrchDatOut.dat where LBFGSOC and MCSRCHOC
 *           write them; WorkArrayOut.dat, DiagOut.dat, IflagOut.dat and
 *           InfoOut.dat (format I2) always.  Moving the *Out.dat files onto
 *           the input names before the next run is the caller's, as with the
 *           reference program.
 *   Text scalars are written with the fewest digits that read back exactly;
 *   records use 4-byte markers (8-byte ones are read too); a W record of
 *   2 GiB or more is SFD_EINVAL.  iflag and info receive IFLAG and MYINFO
 *   as written; x (n, in/out) becomes the next model.  Unreadable or
 *   inconsistent files (another n or m, ONE/ZERO not 1/0) are SFD_EIO. */
int  sfd_lbfgsoc_step_files(const char *dir, const sfd_lbfgsoc_opts *o, int n, double *x, double f,
                            const double *g, int *iflag, int *info);

/* ---- In-process L-BFGS driver (libs/libwave/Inversion.f90:29-189 l_bfgs) --
 * Options mirror InversionParam (Inversion_types.f90:46-58): niter (default
 * 10), neval (3*niter), vpmin/vpmax (XMIN/XMAX), bound (const_type, default 2:
 * clip + FREEZE_X at each accepted iteration; 1: clip at each evaluation),
 * freeze_soft (default 1), MSAVE = m (5), eval0 = invparam%eval at entry. */
typedef struct {
    int    niter, neval, m, const_type, freeze_soft, eval0;
    double xmin, xmax;
} sfd_lbfgs_opts;

/* Run report: iflag = the final IFLAG, info = MCSRCH's last INFO
 * (lbfgs.f:652-700 codes); the history arrays are optional caller buffers of hist_cap
 * entries (one per evaluation): f_eval = unscaled misfit, t_eval = wall
 * seconds since the start, accepted = 1 where the evaluation completed an
 * iteration (MYINFO = 1). */
typedef struct {
    int     iterations, evaluations, iflag, info, nhist, hist_cap;
    double  seconds;
    double *f_eval, *t_eval;
    int    *accepted;
    /* learned preconditioner (sfd_fwi_lbfgs with precond): accepted iterations
     * whose direction used the network, line searches abandoned and restarted
     * plain, directions replaced because they were not descent directions,
     * training wall time, training loss before / after */
    int     pc_iterations, pc_fallbacks, pc_rejects;
    double  pc_train_seconds, pc_loss_first, pc_loss_last;
} sfd_lbfgs_report;

/* Learned gradient preconditioner (no reference counterpart): a 4-layer 3x3
 * CNN (1 -> 16 -> 16 -> 16 -> 1 channels, BF16 weights and activations, FP32
 * accumulation on MFMA) trained on the GPU from the (gradient, accepted step)
 * pairs of the first train_after L-BFGS iterations and then used as the
 * initial inverse Hessian H0 of the two-loop recursion, scaled as the plain
 * H0 is (ys / y.H0(y)).  A preconditioned direction that is not a descent
 * direction is replaced by the plain one; a preconditioned line search that
 * fails or needs more than max_evals evaluations is restarted plain from the
 * same model; after max_fallbacks such restarts the network is switched off. */
typedef struct {
    int   enable;          /* 0: plain L-BFGS                                  */
    int   train_after;     /* accepted iterations whose pairs train the net (3) */
    int   train_steps;     /* Adam steps (300)                                 */
    float lr;              /* Adam learning rate (2e-3)                        */
    int   seed;            /* weight initialisation                            */
    int   max_evals;       /* evaluations allowed per preconditioned search (4);
                              0 abandons every preconditioned search at once */
    int   max_fallbacks;   /* restarts before the network is switched off (3)  */
    int   target;          /* training pairs: 0 gradient g_k -> accepted step
                              s_k of that iteration; 1 g_k -> the accepted model
                              update from x_k to the newest model, sum_{j>=k} s_j;
                              2 the secant pairs y_k = g_{k+1} - g_k -> s_k (the
                              quasi-Newton condition H y = s), in the coordinates
                              of the illumination wrap when illum_h0 is set */
    int   use_net;         /* 1: the network is trained and applied (default)   */
    int   illum_h0;        /* 1: wrap the operator in the illumination diagonal
                              W = illum^(-1/2) (compute_fct_gdt's illumination at
                              the newest evaluated model): H0 = g W N(W q)     */
    int   coarsen;         /* 0 or 1: the network runs on the model grid; r > 1:
                              on the r-times coarser grid (r x r block means R,
                              bilinear prolongation P back): H0 = g W P N(R W q),
                              trained on R g_k -> R s_k; with use_net 0 the
                              smoother g W P R W q alone                        */
} sfd_precond_opts;

/* fctgdt of Inversion.f90:31-39 with the model passed explicitly: evaluate
 * the misfit f and its gradient g (FP32, the reference's sign: LBFGS is given
 * -g) at the FP32 model x; return 0 on success. */
typedef int (*sfd_fctgdt)(void *ctx, const float *x, float *g, double *f);

/* l_bfgs: x (n, FP32) is the starting model on entry and the result on exit
 * (the last accepted model, or the next trial model when niter is reached,
 * as the reference returns).  vpmask / vpinit may be NULL (all ones / x). */
int  sfd_lbfgs_run(const sfd_lbfgs_opts *o, int n, float *x, const float *vpmask,
                   const double *vpinit, sfd_fctgdt fct, void *ctx,
                   sfd_lbfgs_report *rep);

/* FWI with l_bfgs, every evaluation being sfd_fwi_gradient on the engine's
 * model (progs/afwi: compute_fct_gdt as fctgdt).  x: (nz, nx) column-major
 * starting velocity, overwritten with the result. */
int  sfd_fwi_lbfgs(sfd_handle h, const sfd_shot *shots, int nshots, const float *dobs,
                   const float *mute, const float *vpmask, const sfd_fwi_opts *fo,
                   const sfd_lbfgs_opts *lo, const sfd_precond_opts *pc, float *x,
                   sfd_lbfgs_report *rep);

/* The preconditioner's network on its own (for tests and offline use).
 * Parameters: SFD_CNN_NPARAM FP32 values, per layer l = 0..3 the weights
 * W[t][co][ci] (t = 3x3 tap, (dz, dx) = (t%3 - 1, t/3 - 1); 16x16 channels,
 * only ci = 0 used in layer 0 and co = 0 in layer 3) then 16 biases.
 * Images: nimg stacked (nz, nx) column-major FP32 arrays. */
#define SFD_CNN_NPARAM 9280
typedef struct sfd_cnn *sfd_cnn_handle;
int  sfd_cnn_create(sfd_cnn_handle *c, int nz, int nx, int nimg_max, int seed);
int  sfd_cnn_destroy(sfd_cnn_handle c);
int  sfd_cnn_get_params(sfd_cnn_handle c, float *theta);
int  sfd_cnn_set_params(sfd_cnn_handle c, const float *theta);
int  sfd_cnn_forward(sfd_cnn_handle c, int nimg, const float *x, float *y);
/* MSE loss mean((y - target)^2) and its gradient with respect to every parameter */
int  sfd_cnn_loss_grad(sfd_cnn_handle c, int nimg, const float *x, const float *target,
                       double *loss, float *grad);
int  sfd_cnn_train(sfd_cnn_handle c, int nimg, const float *x, const float *target, int steps,
                   float lr, double *loss_first, double *loss_last);

/* ---- 3D forward modeling (genpar%twoD false: Acoustic_modeling_sep.f90:
 *      112-121; constant density, nbound 4).  A handle of its own: model
 *      vel(nz,nx,ny) column-major (z fastest, then x, then y). -------------- */
typedef struct sfd_grid3d {
    int32_t nz, nx, ny;      /* model samples: depth, x, y                     */
    float   oz, ox, oy;      /* origin (m)                                     */
    float   dz, dx, dy;      /* spacing (m)                                    */
    float   dt;              /* time step (s)                                  */
    int32_t ntaper;          /* absorbing pad per side (cells), >= 10          */
    int32_t lsinc;           /* sinc interpolator length (7)                   */
    float   aperture_x, aperture_y;   /* migration apertures (m), ExtractPadModel.f90:55-58 */
    float   fmax;            /* vel_check's frequency (Hz); 0: 30              */
    int32_t snapi;           /* RTM snapshot interval (steps)                  */
} sfd_grid3d;

/* One 3D shot: coord(1..3) = z, x, y of the source and of each receiver. */
typedef struct sfd_shot3d {
    float        sz, sx, sy;
    int32_t      nt, ntr;
    const float *wavelet;
    const float *rz, *rx, *ry;
} sfd_shot3d;

typedef struct sfd_ctx3d *sfd_handle3d;
int  sfd_create3d(sfd_handle3d *h, const sfd_grid3d *g);
int  sfd_destroy3d(sfd_handle3d h);
/* h NULL: the message of the last failed sfd_create3d. */
const char *sfd_last_error3d(sfd_handle3d h);
/* readvel + vel_check in 3D (Courant stop 0.35, ReadSouVelRho.f90:629-675);
 * vel_bwd (may be NULL): RTM's backward-pass model (mod%vel2,
 * to_disk_to_memory_RTM.f90:425). */
int  sfd_set_model3d(sfd_handle3d h, const float *vel, const float *vel_bwd);
/* genpar%withRho in 3D (ExtractPadModel.f90:259-270, Acoustic_modeling_sep.f90:
 * 113-121): rho(nz, nx, ny) as vel, or NULL for constant density.  With a
 * density, sfd_forward3d windows and pads rho as vel, takes 1/rho2 from the
 * 3D Interpolate (Interpolate.f90:44-130), runs
 * FD_3D_derivatives_acoustic_forward_grid (FDderivatives.f90:263-382) with
 * FD_acoustic_rho_init_coefs over delta, injects with the v^2 rho gain
 * (Injection.f90:181-187) and shifts each source wavelet half a time step
 * as readsou does (ReadSouVelRho.f90:328-347). */
int  sfd_set_density3d(sfd_handle3d h, const float *rho);
/* Forward modeling of nshots 3D shots, one after another on the device: per
 * shot the window of its patch (extract_coord_source_receiver_patch with the
 * x and y apertures), FD_2nd_3D_derivatives_scalar_forward_grid
 * (FDderivatives.f90:97-131), Injection_sinc of the source
 * (Injection.f90:274-346), Boundary_3d's bands in the reference's order
 * (Boundary.f90:89-184) and extraction of extract_kind
 * (SFD_EXTRACT_*: Extraction.f90:71-162,217-246 -> :276-337).  As in the
 * reference, the 3D sinc injection and extraction weight the y axis with the
 * x sinc (Injection.f90:314, Extraction.f90:315), and SFD_EXTRACT_SINC_V2
 * scales by v, not v^2, in 3D (Extraction.f90:128).  traces: per shot
 * (nt, ntr) time fastest, concatenated. */
int  sfd_forward3d(sfd_handle3d h, const sfd_shot3d *shots, int nshots, float *traces_out, int extract_kind);
/* 3DRTM.x in 3D (RTM_to_memory, to_disk_to_memory_RTM.f90:358-492, twoD
 * false): per shot forward modeling storing u2 on the window every snapi
 * steps (Extraction_wavefield, Extraction.f90:562-601), then reverse time
 * with the same forward stencil (the 3D branch calls
 * FD_2nd_3D_derivatives_scalar_forward_grid, or the density operator), the
 * backward model when set, Injection_sinc of every receiver's trace in trace
 * order and Imaging_condition_sourceonly_from_memory (Imaging.f90:228-275:
 * image += u S taperx taper_y, illum += S S taperx tapery); the window image
 * is added into image / illum (nz,nx,ny) as mod_copy_image does
 * (ExtractPadModel.f90:662-695, x and y by nearest index).  traces_in: per
 * shot (nt, ntr) time fastest, concatenated. */
int  sfd_rtm3d(sfd_handle3d h, const sfd_shot3d *shots, int nshots, const float *traces_in, float *image,
               float *illum);
/* 3DRTM.x LSRTM=1 in 3D (LSRTM_to_memory, to_disk_to_memory_RTM.f90:494-579,
 * twoD false): the forward pass with the second-differenced source
 * (ReadSouVelRho.f90:318-326) storing snapshots, then reverse time on the
 * backward model when set with the adjoint-style 3D stencil
 * (FDderivatives.f90:780-819), Injection_LSRTM_sinc of the traces (weight
 * 1/(dz dx dy), Injection.f90:239-272) and
 * Imaging_condition_LSRTM_sourceonly_from_memory (Imaging.f90:698-745);
 * image / illum (nz,nx,ny) accumulate as in sfd_rtm3d.  Constant density. */
int  sfd_lsrtm3d(sfd_handle3d h, const sfd_shot3d *shots, int nshots, const float *traces_in, float *image,
                 float *illum);
/* 3DBORNMOD.x in 3D (BORN_to_memory, to_disk_to_memory_BORN.f90:123-220,
 * twoD false): the same forward pass, then a second forward pass on the
 * backward model when set with no source, Injection_Born (Imaging.f90:747-788,
 * 2 taper m S / v added to u3 every snapi steps) and
 * Extraction_array_LSRTM_sinc.  refl (nz,nx,ny) is windowed like vel.
 * traces_out: per shot (nt, ntr), concatenated.  Constant density. */
int  sfd_born3d(sfd_handle3d h, const sfd_shot3d *shots, int nshots, const float *refl, float *traces_out);
/* 3DAFWI.x's gradient in 3D (AFWI_to_memory, to_disk_to_memory_AFWI.f90:
 * 229-415, twoD false): as sfd_afwi_gradient, with the 3D branch's forward
 * stencil in the reverse pass (:394-403) and, with sfd_set_density3d, the
 * v-rho imaging condition with its y term (Imaging.f90:424-512, dk = 1
 * unless the window has one line).  dobs / mute: per shot (nt, ntr)
 * concatenated (mute may be NULL); grad: (nz,nx,ny) vp [, (nz,nx,ny) rho or
 * Ip], overwritten with the sum over the shots; illum may be NULL.  FP32
 * snapshots only (opt->snap_mode SFD_SNAP_FP32). */
int  sfd_afwi_gradient3d(sfd_handle3d h, const sfd_shot3d *shots, int nshots, const float *dobs, const float *mute,
                         const sfd_afwi_opts *opt, double *misfit, float *grad, float *illum);

/* ---- SEP files (the reference reads and writes every array through SEPlib:
 *      from_aux / sreed / srite, libs/libwave/IO.f90:14-51,
 *      ReadSouVelRho.f90:59-297).  Host-side only: these calls touch no GPU.
 * A history file holds name=value assignments (the last one of a name wins;
 * values may be quoted); `in` names the binary file, `esize` must be 4,
 * `data_format` is "xdr_float" (big-endian, the default when absent) or
 * "native_float". ------------------------------------------------------- */
typedef struct sfd_sep_header {
    int32_t ndim;            /* highest axis with n > 1 (>= 1)                 */
    int64_t n[6];            /* n1..n6 (1 when absent)                         */
    float   o[6], d[6];      /* o1..o6 (0), d1..d6 (1)                          */
    int32_t esize;           /* bytes per element: 4                           */
    int32_t xdr;             /* 1: big-endian data (xdr_float), 0: native       */
    char    in[1024];        /* the binary file (resolved against the header's
                                directory when not found as given)             */
} sfd_sep_header;

/* from_aux(tag, 'n1', …): the header of a history file. */
int  sfd_sep_read_header(const char *hfile, sfd_sep_header *hdr);
/* sreed: count floats from element offset of the file's data, converted to
 * native byte order. */
int  sfd_sep_read(const char *hfile, size_t offset, size_t count, float *data);
/* srite + to_history: writes count = n1*…*n6 floats (native, to hdr->in or,
 * when that is empty, to hfile followed by '@') and the history file. */
int  sfd_sep_write(const char *hfile, const sfd_sep_header *hdr, const float *data, size_t count);
/* readgathercoords (ReadSouVelRho.f90:161-241) with number_of_shots /
 * number_of_traces (:828-890): the coordinate file holds n1 keys per trace
 * and n2 traces; keys are the 1-based indexes keygx, keygy, keysx, keysy,
 * keyselev, keygelev.  A shot is a run of consecutive traces with equal
 * (sx, sy); its source (selev, sx) is read from its last trace, each trace's
 * receiver is (gelev, gx).  Call with shot_first NULL to get *nshots only;
 * then with arrays of cap_shots shots (shot_first, shot_ntr, sz, sx) and
 * cap_traces traces (rz, rx); any of sz, sx, rz, rx may be NULL. */
int  sfd_sep_gathers(const char *coordfile, const int keys[6], int *nshots, int *shot_first, int *shot_ntr,
                     float *sz, float *sx, float *rz, float *rx, int cap_shots, int cap_traces);

#ifdef __cplusplus
}
#endif
#endif /* SEISFD_H */
