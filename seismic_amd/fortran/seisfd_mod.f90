! seisfd_mod.f90 -- Fortran 2003 binding of the seisfd C-ABI (include/seisfd.h).
!
! A Fortran host program of the reference (libs/libwave, progs/rtm, progs/afwi)
! uses this module instead of calling propagator_acoustic
! (libs/libwave/Propagator.f90:27-270) or l_bfgs (libs/libwave/Inversion.f90:29-189):
! the derived types below are bind(C) twins of the C structs, and every entry
! point is an interface to the extern "C" symbol of libseisfd.so.  Arrays are
! passed as they are held: vel(nz,nx), trace(nt,ntr), column-major.
module seisfd_mod
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: SFD_OK = 0, SFD_EINVAL = -1, SFD_ENOMEM = -2, SFD_EHIP = -3, &
       SFD_EUNSUPPORTED = -4, SFD_ESTATE = -5, SFD_EUNSTABLE = -6, SFD_ENOTFINITE = -7, SFD_EIO = -8, &
       SFD_EALIASED = -9
  integer(c_int), parameter :: SFD_SNAP_FP32 = 0, SFD_SNAP_FP16S = 1, SFD_SNAP_FP16G = 2
  integer(c_int), parameter :: SFD_NORM_L2 = 2, SFD_NORM_L1 = 1, SFD_NORM_HUBER = 12, SFD_NORM_CAUCHY = 3, &
       SFD_NORM_HYPERBOLIC = 4
  integer(c_int), parameter :: SFD_EXTRACT_SINC_V2 = 0, SFD_EXTRACT_NEAREST = 1, &
       SFD_EXTRACT_SINC_LSRTM = 2, SFD_EXTRACT_SINC_UNIT = 3

  type, bind(C) :: sfd_grid
     integer(c_int32_t) :: nz, nx
     real(c_float)      :: oz, ox, dz, dx, dt
     integer(c_int32_t) :: ntaper, lsinc, snapi
     real(c_float)      :: aperture_x
     integer(c_int32_t) :: rec_type, shot_type, surf_type
     real(c_float)      :: fmax = 0.   ! vel_check's fmax (Hz); 0: 30
  end type sfd_grid

  type, bind(C) :: sfd_shot
     real(c_float)      :: sz, sx
     integer(c_int32_t) :: nt, ntr
     type(c_ptr)        :: wavelet, rz, rx
  end type sfd_shot

  ! 3D forward modeling (sfd_create3d ...): model vel(nz, nx, ny)
  type, bind(C) :: sfd_grid3d
     integer(c_int32_t) :: nz, nx, ny
     real(c_float)      :: oz, ox, oy, dz, dx, dy, dt
     integer(c_int32_t) :: ntaper, lsinc
     real(c_float)      :: aperture_x, aperture_y, fmax
     integer(c_int32_t) :: snapi = 1
  end type sfd_grid3d

  type, bind(C) :: sfd_shot3d
     real(c_float)      :: sz, sx, sy
     integer(c_int32_t) :: nt, ntr
     type(c_ptr)        :: wavelet, rz, rx, ry
  end type sfd_shot3d

  type, bind(C) :: sfd_fwi_opts
     real(c_float)      :: illupow
     integer(c_int32_t) :: rect_z, rect_x, do_z, do_x, snap_mode
     integer(c_int32_t) :: norm = 0          ! SFD_NORM_* (0 = L2), data_nrm_type
     real(c_float)      :: norm_thresh = 0.  ! data_threshold
  end type sfd_fwi_opts

  type, bind(C) :: sfd_afwi_opts
     integer(c_int32_t) :: snap_mode = 0, norm = 0
     real(c_float)      :: norm_thresh = 0.
     integer(c_int32_t) :: vprho_param = 0
  end type sfd_afwi_opts

  type, bind(C) :: sfd_bandpass_opts
     real(c_float)      :: flo = 0., fhi = 20.
     integer(c_int32_t) :: nplo = 14, nphi = 14, phase = 0
  end type sfd_bandpass_opts

  type, bind(C) :: sfd_mute_opts
     real(c_float) :: v0 = 1500., t0 = 0., v1 = 10000., t1 = 40., tramp = 0.5
  end type sfd_mute_opts

  ! SEP history-file header (sfd_sep_read_header / sfd_sep_write)
  type, bind(C) :: sfd_sep_header
     integer(c_int32_t) :: ndim
     integer(c_int64_t) :: n(6)
     real(c_float)      :: o(6), d(6)
     integer(c_int32_t) :: esize, xdr
     character(kind=c_char) :: in(1024)
  end type sfd_sep_header

  type, bind(C) :: sfd_lbfgs_opts
     integer(c_int) :: niter, neval, m, const_type, freeze_soft, eval0
     real(c_double) :: xmin, xmax
  end type sfd_lbfgs_opts

  ! out-of-core L-BFGS options (LBFGSOC of libs/libnlinv/nlinv_lbfgs.f)
  type, bind(C) :: sfd_lbfgsoc_opts
     integer(c_int) :: m, nparam
     type(c_ptr)    :: xmin, xmax      ! c_loc of nparam real(c_double) bounds
     integer(c_int) :: stp1_opt
     real(c_double) :: step_init
     integer(c_int) :: clip_type
  end type sfd_lbfgsoc_opts

  type, bind(C) :: sfd_lbfgs_report
     integer(c_int) :: iterations, evaluations, iflag, info, nhist, hist_cap
     real(c_double) :: seconds
     type(c_ptr)    :: f_eval, t_eval, accepted
     integer(c_int) :: pc_iterations, pc_fallbacks, pc_rejects
     real(c_double) :: pc_train_seconds, pc_loss_first, pc_loss_last
  end type sfd_lbfgs_report

  type, bind(C) :: sfd_precond_opts
     integer(c_int) :: enable, train_after, train_steps
     real(c_float)  :: lr
     integer(c_int) :: seed, max_evals, max_fallbacks, target, use_net, illum_h0, coarsen
  end type sfd_precond_opts

  interface
     integer(c_int) function sfd_create(h, g) bind(C, name='sfd_create')
       import :: c_int, c_ptr, sfd_grid
       type(c_ptr), intent(out) :: h
       type(sfd_grid), intent(in) :: g
     end function sfd_create

     integer(c_int) function sfd_destroy(h) bind(C, name='sfd_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function sfd_destroy

     type(c_ptr) function sfd_last_error(h) bind(C, name='sfd_last_error')
       import :: c_ptr
       type(c_ptr), value :: h
     end function sfd_last_error

     integer(c_int) function sfd_set_model(h, vel, vel_bwd) bind(C, name='sfd_set_model')
       import :: c_int, c_ptr, c_float
       type(c_ptr), value :: h
       real(c_float), intent(in) :: vel(*)
       type(c_ptr), value :: vel_bwd
     end function sfd_set_model

     integer(c_int) function sfd_set_density(h, rho) bind(C, name='sfd_set_density')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       type(c_ptr), value :: rho
     end function sfd_set_density

     integer(c_int) function sfd_set_elevation(h, elev) bind(C, name='sfd_set_elevation')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       type(c_ptr), value :: elev
     end function sfd_set_elevation

     integer(c_int) function sfd_create3d(h, g) bind(C, name='sfd_create3d')
       import :: c_int, c_ptr, sfd_grid3d
       type(c_ptr), intent(out) :: h
       type(sfd_grid3d), intent(in) :: g
     end function sfd_create3d

     integer(c_int) function sfd_destroy3d(h) bind(C, name='sfd_destroy3d')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function sfd_destroy3d

     type(c_ptr) function sfd_last_error3d(h) bind(C, name='sfd_last_error3d')
       import :: c_ptr
       type(c_ptr), value :: h
     end function sfd_last_error3d

     integer(c_int) function sfd_set_model3d(h, vel, vel_bwd) bind(C, name='sfd_set_model3d')
       import :: c_int, c_ptr, c_float
       type(c_ptr), value :: h
       real(c_float), intent(in) :: vel(*)
       type(c_ptr), value :: vel_bwd   ! c_loc of the backward model, or c_null_ptr
     end function sfd_set_model3d

     integer(c_int) function sfd_rtm3d(h, shots, nshots, traces, image, illum) bind(C, name='sfd_rtm3d')
       import :: c_int, c_ptr, c_float, sfd_shot3d
       type(c_ptr), value :: h
       type(sfd_shot3d), intent(in) :: shots(*)
       integer(c_int), value :: nshots
       real(c_float), intent(in) :: traces(*)
       real(c_float), intent(inout) :: image(*), illum(*)   ! (nz, nx, ny), accumulated
     end function sfd_rtm3d

     integer(c_int) function sfd_lsrtm3d(h, shots, nshots, traces, image, illum) bind(C, name='sfd_lsrtm3d')
       import :: c_int, c_ptr, c_float, sfd_shot3d
       type(c_ptr), value :: h
       type(sfd_shot3d), intent(in) :: shots(*)
       integer(c_int), value :: nshots
       real(c_float), intent(in) :: traces(*)
       real(c_float), intent(inout) :: image(*), illum(*)   ! (nz, nx, ny), accumulated
     end function sfd_lsrtm3d

     integer(c_int) function sfd_born3d(h, shots, nshots, refl, traces) bind(C, name='sfd_born3d')
       import :: c_int, c_ptr, c_float, sfd_shot3d
       type(c_ptr), value :: h
       type(sfd_shot3d), intent(in) :: shots(*)
       integer(c_int), value :: nshots
       real(c_float), intent(in) :: refl(*)                ! (nz, nx, ny)
       real(c_float), intent(inout) :: traces(*)
     end function sfd_born3d

     integer(c_int) function sfd_afwi_gradient3d(h, shots, nshots, dobs, mute, opt, misfit, grad, illum) &
          bind(C, name='sfd_afwi_gradient3d')
       import :: c_int, c_ptr, c_float, c_double, sfd_shot3d, sfd_afwi_opts
       type(c_ptr), value :: h
       type(sfd_shot3d), intent(in) :: shots(*)
       integer(c_int), value :: nshots
       real(c_float), intent(in) :: dobs(*)
       type(c_ptr), value :: mute                ! c_loc of the mute, or c_null_ptr
       type(sfd_afwi_opts), intent(in) :: opt
       real(c_double), intent(out) :: misfit
       real(c_float), intent(inout) :: grad(*)   ! (nz,nx,ny) vp [, (nz,nx,ny) rho or Ip]
       type(c_ptr), value :: illum
     end function sfd_afwi_gradient3d

     integer(c_int) function sfd_set_density3d(h, rho) bind(C, name='sfd_set_density3d')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       type(c_ptr), value :: rho     ! c_loc of rho(nz, nx, ny), or c_null_ptr
     end function sfd_set_density3d

     integer(c_int) function sfd_forward3d(h, shots, nshots, traces, extract_kind) bind(C, name='sfd_forward3d')
       import :: c_int, c_ptr, c_float, sfd_shot3d
       type(c_ptr), value :: h
       type(sfd_shot3d), intent(in) :: shots(*)
       integer(c_int), value :: nshots, extract_kind
       real(c_float), intent(inout) :: traces(*)
     end function sfd_forward3d

     integer(c_int) function sfd_afwi_gradient(h, shots, nshots, dobs, mute, opt, misfit, grad, illum) &
          bind(C, name='sfd_afwi_gradient')
       import :: c_int, c_ptr, c_float, c_double, sfd_shot, sfd_afwi_opts
       type(c_ptr), value :: h
       type(sfd_shot), intent(in) :: shots(*)
       integer(c_int), value :: nshots
       type(c_ptr), value :: dobs, mute          ! c_loc of the arrays, or c_null_ptr
       type(sfd_afwi_opts), intent(in) :: opt
       real(c_double), intent(out) :: misfit
       real(c_float), intent(inout) :: grad(*)   ! (nz,nx) vp [, (nz,nx) rho or Ip]
       type(c_ptr), value :: illum
     end function sfd_afwi_gradient

     integer(c_int) function sfd_forward(h, shots, nshots, traces, extract_kind) bind(C, name='sfd_forward')
       import :: c_int, c_ptr, c_float, sfd_shot
       type(c_ptr), value :: h
       type(sfd_shot), intent(in) :: shots(*)
       integer(c_int), value :: nshots, extract_kind
       real(c_float), intent(inout) :: traces(*)
     end function sfd_forward

     integer(c_int) function sfd_rtm(h, shots, nshots, traces, image, illum, snap_mode) bind(C, name='sfd_rtm')
       import :: c_int, c_ptr, c_float, sfd_shot
       type(c_ptr), value :: h
       type(sfd_shot), intent(in) :: shots(*)
       integer(c_int), value :: nshots, snap_mode
       real(c_float), intent(in) :: traces(*)
       real(c_float), intent(inout) :: image(*), illum(*)
     end function sfd_rtm

     ! keep traces in HBM for sfd_rtm / sfd_lsrtm called with a NULL traces_in
     ! (bind sfd_rtm with a type(c_ptr), value argument to pass c_null_ptr)
     integer(c_int) function sfd_set_traces(h, traces, count) bind(C, name='sfd_set_traces')
       import :: c_int, c_ptr, c_float, c_size_t
       type(c_ptr), value :: h
       real(c_float), intent(in) :: traces(*)
       integer(c_size_t), value :: count
     end function sfd_set_traces

     integer(c_int) function sfd_born(h, shots, nshots, refl, traces) bind(C, name='sfd_born')
       import :: c_int, c_ptr, c_float, sfd_shot
       type(c_ptr), value :: h
       type(sfd_shot), intent(in) :: shots(*)
       integer(c_int), value :: nshots
       real(c_float), intent(in) :: refl(*)
       real(c_float), intent(inout) :: traces(*)
     end function sfd_born

     integer(c_int) function sfd_adjoint_exact(h, shots, nshots, dbar, sbar, extract_kind) &
          bind(C, name='sfd_adjoint_exact')
       import :: c_int, c_ptr, c_float, sfd_shot
       type(c_ptr), value :: h
       type(sfd_shot), intent(in) :: shots(*)
       integer(c_int), value :: nshots, extract_kind
       real(c_float), intent(in) :: dbar(*)
       real(c_float), intent(inout) :: sbar(*)
     end function sfd_adjoint_exact

     integer(c_int) function sfd_born_adjoint_exact(h, shots, nshots, dbar, mbar) &
          bind(C, name='sfd_born_adjoint_exact')
       import :: c_int, c_ptr, c_float, sfd_shot
       type(c_ptr), value :: h
       type(sfd_shot), intent(in) :: shots(*)
       integer(c_int), value :: nshots
       real(c_float), intent(in) :: dbar(*)
       real(c_float), intent(inout) :: mbar(*)
     end function sfd_born_adjoint_exact

     integer(c_int) function sfd_lsrtm(h, shots, nshots, traces, image, illum) bind(C, name='sfd_lsrtm')
       import :: c_int, c_ptr, c_float, sfd_shot
       type(c_ptr), value :: h
       type(sfd_shot), intent(in) :: shots(*)
       integer(c_int), value :: nshots
       real(c_float), intent(in) :: traces(*)
       real(c_float), intent(inout) :: image(*), illum(*)
     end function sfd_lsrtm

     ! SEP files (host side): from_aux / sreed / srite of the reference
     integer(c_int) function sfd_sep_read_header(hfile, hdr) bind(C, name='sfd_sep_read_header')
       import :: c_int, c_char, sfd_sep_header
       character(kind=c_char), intent(in) :: hfile(*)
       type(sfd_sep_header), intent(out) :: hdr
     end function sfd_sep_read_header

     integer(c_int) function sfd_sep_read(hfile, offset, count, data) bind(C, name='sfd_sep_read')
       import :: c_int, c_char, c_size_t, c_float
       character(kind=c_char), intent(in) :: hfile(*)
       integer(c_size_t), value :: offset, count
       real(c_float), intent(out) :: data(*)
     end function sfd_sep_read

     integer(c_int) function sfd_sep_write(hfile, hdr, data, count) bind(C, name='sfd_sep_write')
       import :: c_int, c_char, c_size_t, c_float, sfd_sep_header
       character(kind=c_char), intent(in) :: hfile(*)
       type(sfd_sep_header), intent(in) :: hdr
       real(c_float), intent(in) :: data(*)
       integer(c_size_t), value :: count
     end function sfd_sep_write

     integer(c_int) function sfd_sep_gathers(coordfile, keys, nshots, shot_first, shot_ntr, sz, sx, rz, rx, &
          cap_shots, cap_traces) bind(C, name='sfd_sep_gathers')
       import :: c_int, c_char, c_ptr
       character(kind=c_char), intent(in) :: coordfile(*)
       integer(c_int), intent(in) :: keys(6)
       integer(c_int), intent(out) :: nshots
       type(c_ptr), value :: shot_first, shot_ntr, sz, sx, rz, rx
       integer(c_int), value :: cap_shots, cap_traces
     end function sfd_sep_gathers

     integer(c_int) function sfd_bandpass(h, opt, nt, ntr, traces) bind(C, name='sfd_bandpass')
       import :: c_int, c_ptr, c_float, sfd_bandpass_opts
       type(c_ptr), value :: h
       type(sfd_bandpass_opts), intent(in) :: opt
       integer(c_int), value :: nt, ntr
       real(c_float), intent(inout) :: traces(*)
     end function sfd_bandpass

     integer(c_int) function sfd_mute_mask(h, shots, nshots, ot, opt, mask) bind(C, name='sfd_mute_mask')
       import :: c_int, c_ptr, c_float, sfd_shot, sfd_mute_opts
       type(c_ptr), value :: h
       type(sfd_shot), intent(in) :: shots(*)
       integer(c_int), value :: nshots
       real(c_float), value :: ot
       type(sfd_mute_opts), intent(in) :: opt
       real(c_float), intent(inout) :: mask(*)
     end function sfd_mute_mask

     integer(c_int) function sfd_fwi_gradient(h, shots, nshots, dobs, mute, vpmask, opt, misfit, grad, &
          illum, resid) bind(C, name='sfd_fwi_gradient')
       import :: c_int, c_ptr, c_float, c_double, sfd_shot, sfd_fwi_opts
       type(c_ptr), value :: h
       type(sfd_shot), intent(in) :: shots(*)
       integer(c_int), value :: nshots
       real(c_float), intent(in) :: dobs(*)
       type(c_ptr), value :: mute, vpmask
       type(sfd_fwi_opts), intent(in) :: opt
       real(c_double), intent(out) :: misfit
       real(c_float), intent(inout) :: grad(*)
       type(c_ptr), value :: illum, resid
     end function sfd_fwi_gradient

     ! one LBFGSOC call; state_in / state_out are NUL-terminated file names
     integer(c_int) function sfd_lbfgsoc_step(state_in, state_out, o, n, x, f, g, iflag, info) &
          bind(C, name='sfd_lbfgsoc_step')
       import :: c_int, c_char, c_double, sfd_lbfgsoc_opts
       character(kind=c_char), intent(in) :: state_in(*), state_out(*)
       type(sfd_lbfgsoc_opts), intent(in) :: o
       integer(c_int), value :: n
       real(c_double), intent(inout) :: x(*)
       real(c_double), value :: f
       real(c_double), intent(in) :: g(*)
       integer(c_int), intent(inout) :: iflag, info
     end function sfd_lbfgsoc_step

     ! one run of progs/inversion/L_BFGS.f90 on the reference's state files in dir
     integer(c_int) function sfd_lbfgsoc_step_files(dir, o, n, x, f, g, iflag, info) &
          bind(C, name='sfd_lbfgsoc_step_files')
       import :: c_int, c_char, c_double, sfd_lbfgsoc_opts
       character(kind=c_char), intent(in) :: dir(*)
       type(sfd_lbfgsoc_opts), intent(in) :: o
       integer(c_int), value :: n
       real(c_double), intent(inout) :: x(*)
       real(c_double), value :: f
       real(c_double), intent(in) :: g(*)
       integer(c_int), intent(out) :: iflag, info
     end function sfd_lbfgsoc_step_files

     integer(c_int) function sfd_lbfgs_run(o, n, x, vpmask, vpinit, fct, ctx, rep) bind(C, name='sfd_lbfgs_run')
       import :: c_int, c_ptr, c_float, c_funptr, sfd_lbfgs_opts, sfd_lbfgs_report
       type(sfd_lbfgs_opts), intent(in) :: o
       integer(c_int), value :: n
       real(c_float), intent(inout) :: x(*)
       type(c_ptr), value :: vpmask, vpinit
       type(c_funptr), value :: fct
       type(c_ptr), value :: ctx
       type(sfd_lbfgs_report), intent(inout) :: rep
     end function sfd_lbfgs_run

     integer(c_int) function sfd_fwi_lbfgs(h, shots, nshots, dobs, mute, vpmask, fo, lo, pc, x, rep) &
          bind(C, name='sfd_fwi_lbfgs')
       import :: c_int, c_ptr, c_float, sfd_shot, sfd_fwi_opts, sfd_lbfgs_opts, sfd_lbfgs_report
       type(c_ptr), value :: h
       type(sfd_shot), intent(in) :: shots(*)
       integer(c_int), value :: nshots
       real(c_float), intent(in) :: dobs(*)
       type(c_ptr), value :: mute, vpmask
       type(sfd_fwi_opts), intent(in) :: fo
       type(sfd_lbfgs_opts), intent(in) :: lo
       type(c_ptr), value :: pc     ! c_loc of a sfd_precond_opts, or c_null_ptr for plain L-BFGS
       real(c_float), intent(inout) :: x(*)
       type(sfd_lbfgs_report), intent(inout) :: rep
     end function sfd_fwi_lbfgs
  end interface

contains

  ! sfd_last_error as a Fortran string
  function sfd_error_message(h) result(msg)
    type(c_ptr), intent(in) :: h
    character(len=256) :: msg
    character(kind=c_char), pointer :: s(:)
    integer :: i
    msg = ' '
    if (.not. c_associated(sfd_last_error(h))) return
    call c_f_pointer(sfd_last_error(h), s, [256])
    do i = 1, 256
       if (s(i) == c_null_char) exit
       msg(i:i) = s(i)
    end do
  end function sfd_error_message

end module seisfd_mod
