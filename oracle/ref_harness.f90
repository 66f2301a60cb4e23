! ref_harness.f90 — TEST INFRASTRUCTURE ONLY.
!
! C-callable harness around the reference's own SEP-free Fortran kernels
! (compiled in place from /root/reference by oracle/Makefile into
! oracle/_ref/libseisref.so).  It composes those kernels in the order of
! propagator_acoustic (libs/libwave/Propagator.f90:142-243), which itself
! cannot be compiled here: it uses Imaging_mod, which needs SEPlib
! (libs/libwave/Imaging.f90:8).  The harness adds no arithmetic of its own:
! every floating-point operation on the wavefield is a call into the
! reference's modules, and the misfit norms are Norm_mod's own functions.
! Used only to pin oracle/seisoracle.c.
module ref_harness
  use iso_c_binding
  use GeneralParam_types
  use DataSpace_types
  use ModelSpace_types
  use FD_types
  use FDcoefs_assign
  use FD_derivatives
  use FDswaptime_mod
  use Interpolate_mod
  use Boundary_mod
  use Injection_mod
  use Extraction_mod
  use Taper_mod
  use Norm_mod
  implicit none
contains

  ! One trace through the reference's misfit and adjoint-source norms
  ! (libs/libinv/Norm.f90:12-141), with THRESH passed by reference as
  ! Compute_OF_RES_ADJ passes its local copy (the routines may overwrite it).
  function ref_fct_compute(ntype, rd, nd, thresh) result(f) bind(C, name='ref_fct_compute')
    integer(c_int), value :: ntype, nd
    real(c_float), intent(in) :: rd(nd)
    real(c_float), intent(inout) :: thresh
    real(c_double) :: f
    real, allocatable :: r(:)
    real :: th
    allocate(r(nd))
    r = rd
    th = thresh
    f = fct_compute(ntype, r, nd, th)
    thresh = th
    deallocate(r)
  end function ref_fct_compute

  subroutine ref_gdt_compute(ntype, rd, nd, thresh) bind(C, name='ref_gdt_compute')
    integer(c_int), value :: ntype, nd
    real(c_float), intent(inout) :: rd(nd)
    real(c_float), intent(inout) :: thresh
    integer :: stat
    real :: th
    th = thresh
    stat = gdt_compute(ntype, rd, nd, th)
    thresh = th
  end subroutine ref_gdt_compute

  subroutine ref_mksinc(sinc, lsinc, shift) bind(C, name='ref_mksinc')
    integer(c_int), value :: lsinc
    real(c_float), value  :: shift
    real(c_float)         :: sinc(lsinc)
    call mksinc(sinc, lsinc, shift)
  end subroutine ref_mksinc

  subroutine ref_higdon_coeff(nu, coeff) bind(C, name='ref_higdon_coeff')
    real(c_float), value :: nu
    real(c_float)        :: coeff(8)
    call Higdon_coeff(nu, 1.0, 0.7, 0.5, coeff)
  end subroutine ref_higdon_coeff

  subroutine ref_taper(n, t) bind(C, name='ref_taper')
    integer(c_int), value :: n
    real(c_float)         :: t(n)
    type(ModelSpace) :: m
    m%nxw = n
    m%nyw = 1
    call compute_taper(m)
    t = m%taperx
    call compute_taper_close(m)
  end subroutine ref_taper

  ! One propagation in Propagator.f90 order.
  !   stencil: 0 forward_grid (FDderivatives.f90:715), 1 adjoint_grid (:847)
  !   inj    : 0 none, 1 Injection_sinc, 2 Injection_simple_noomp, 3 Injection_LSRTM_sinc
  !   ext    : -1 none, 0 Extraction_array_sinc, 1 Extraction_array_simple_afwi_noomp,
  !            2 Extraction_array_LSRTM_sinc, 3 Extraction_array_sinc_afwi
  !   tdir   : +1 it=1..nt, -1 it=nt..1
  !   wave   : (nz,nxw,nsnap) captured u2 every snapi steps (Extraction_wavefield)
  !   rec_type, shot_type, surf_type: GeneralParam's ghost / free-surface
  !            switches; elevj: elev%elev over the FD columns nmin2..nmax2
  !            (read when surf_type /= 0).  The FD bounds follow
  !            read_window_vel's rule (ExtractPadModel.f90:228-233, which needs
  !            SEPlib and is not compiled); vel_pad must have those bounds.
  !   with_rho: withRho, rho_pad the padded density (as vel_pad): Interpolate,
  !            allocate_sxx2d_szz2d_delp2d, FD_acoustic_rho_init_coefs with
  !            coefpower 1 and FD_2D_derivatives_acoustic_forward_grid as the
  !            stencil whatever `stencil` says, as the RTM and modeling
  !            drivers call them (to_disk_to_memory_RTM.f90:389-397,454-462)
  subroutine ref_propagate(nz, nxw, ntaper, dz, dx, dt, oz, oxw, lsinc, snapi, vel_pad, &
       &                   stencil, inj, nsou, souz, soux, soutr, ext, ndat, datz, datx, dattr, &
       &                   nt, tdir, nsnap, wave, rec_type, shot_type, surf_type, elevj, with_rho, rho_pad) &
       &                   bind(C, name='ref_propagate')
    integer(c_int), value :: nz, nxw, ntaper, lsinc, snapi, stencil, inj, nsou, ext, ndat, nt, tdir, nsnap
    integer(c_int), value :: rec_type, shot_type, surf_type, with_rho
    real(c_float), value  :: dz, dx, dt, oz, oxw
    real(c_float)         :: vel_pad(*), souz(*), soux(*), soutr(nt, *), datz(*), datx(*), dattr(nt, *)
    real(c_float)         :: wave(nz, nxw, *), elevj(*), rho_pad(*)

    type(GeneralParam)         :: genpar
    type(FDbounds)             :: bounds
    type(ModelSpace)           :: model
    type(ModelSpace_elevation) :: elev
    type(HigdonParam)          :: hig
    type(UnscaledFDcoefs)      :: fdcoefs
    type(ScaledFDcoefs), target :: scaled
    type(USpace)               :: grid
    type(WaveSpace)            :: wfld
    type(TraceSpace), allocatable :: sou(:), dat(:)
    integer :: i, j, it, n1, n2

    genpar%twoD = .true.
    genpar%withRho = with_rho /= 0
    genpar%Born = .false.
    genpar%LSRTM = .false.
    genpar%verbose = .false.
    genpar%optim = .true.
    genpar%dt = dt
    genpar%nt = nt
    genpar%delta = (/ dz, dx, 1.0 /)
    genpar%omodel = (/ oz, oxw, 0.0 /)
    genpar%lsinc = lsinc
    genpar%nbound = 0
    genpar%ntaper = ntaper
    genpar%snapi = snapi
    genpar%ntsnap = nsnap
    genpar%rec_type = rec_type
    genpar%shot_type = shot_type
    genpar%surf_type = surf_type
    genpar%coefpower = 2
    if (with_rho /= 0) genpar%coefpower = 1
    if (tdir > 0) then
       genpar%tmin = 1; genpar%tmax = nt; genpar%tstep = 1
    else
       genpar%tmin = nt; genpar%tmax = 1; genpar%tstep = -1
    end if

    if (shot_type > 0 .or. surf_type > 0) then
       if (shot_type > 0) bounds%nmin1 = -(ntaper + lsinc/2 + 2) + 1
       if (surf_type > 0) bounds%nmin1 = -lsinc + 1
    else
       bounds%nmin1 = -ntaper + 1
    end if
    bounds%nmax1 = nz + ntaper
    bounds%nmin2 = -ntaper + 1
    bounds%nmax2 = nxw + ntaper
    bounds%nmin3 = 1
    bounds%nmax3 = 1

    model%nz = nz
    model%nxw = nxw
    model%nyw = 1
    model%dz = dz
    model%dx = dx
    model%dy = 1.0
    n1 = bounds%nmax1 - bounds%nmin1 + 9
    n2 = bounds%nmax2 - bounds%nmin2 + 9
    allocate(model%vel(bounds%nmin1-4:bounds%nmax1+4, bounds%nmin2-4:bounds%nmax2+4, 1:1))
    do j = 1, n2
       do i = 1, n1
          model%vel(bounds%nmin1-5+i, bounds%nmin2-5+j, 1) = vel_pad(i + (j-1)*n1)
       end do
    end do

    if (with_rho /= 0) then
       allocate(model%rho(bounds%nmin1-4:bounds%nmax1+4, bounds%nmin2-4:bounds%nmax2+4, 1:1))
       allocate(model%rho2(bounds%nmin1-4:bounds%nmax1+4, bounds%nmin2-4:bounds%nmax2+4, 1:1))
       do j = 1, n2
          do i = 1, n1
             model%rho(bounds%nmin1-5+i, bounds%nmin2-5+j, 1) = rho_pad(i + (j-1)*n1)
          end do
       end do
       call Interpolate(model, bounds, genpar)
    end if

    allocate(sou(max(nsou,1)), dat(max(ndat,1)))
    do i = 1, nsou
       sou(i)%coord = (/ souz(i), soux(i), 0.0 /)
       allocate(sou(i)%trace(nt, 1))
       sou(i)%trace(:, 1) = soutr(:, i)
    end do
    do i = 1, ndat
       dat(i)%coord = (/ datz(i), datx(i), 0.0 /)
       allocate(dat(i)%trace(nt, 1))
       dat(i)%trace(:, 1) = dattr(:, i)
    end do
    if (nsnap > 0) then
       allocate(wfld%wave(nz, nxw, 1, nsnap, 1))
       wfld%wave = 0.
    end if

    ! Propagator.f90:144-166
    call allocateUSpace(grid, genpar, bounds)
    call zeroUSpace(grid)
    allocate(hig%gx(16, bounds%nmin1:bounds%nmax1, bounds%nmin3:bounds%nmax3))
    allocate(hig%gy(16, bounds%nmin1:bounds%nmax1, bounds%nmin2:bounds%nmax2))
    allocate(hig%gz(16, bounds%nmin2:bounds%nmax2, bounds%nmin3:bounds%nmax3))
    if (with_rho /= 0) then
       call FD_acoustic_rho_init_coefs(fdcoefs)
       call allocate_sxx2d_szz2d_delp2d(bounds, model)
    else
       call FD_acoustic_init_coefs(fdcoefs)
    end if
    call FD_types_assign_scaled_coefs(fdcoefs, scaled, genpar)
    call FD_derivatives_coef_init(scaled)
    call Higdon(genpar%dt, model, bounds, hig)
    if (nsou > 0) call ModelSpace_compute_array_xyz_positions(genpar, sou(1:nsou))
    if (ndat > 0) call ModelSpace_compute_array_xyz_positions(genpar, dat(1:ndat))
    ! Propagator.f90:146-150 allocates the elevation tables; the drivers fill elev
    allocate(elev%elev(bounds%nmin2:bounds%nmax2, 1:1))
    allocate(elev%ielev_z(bounds%nmin2:bounds%nmax2, 1:1))
    allocate(elev%delev_z(bounds%nmin2:bounds%nmax2, 1:1))
    do j = bounds%nmin2, bounds%nmax2
       elev%elev(j, 1) = elevj(j - bounds%nmin2 + 1)
    end do
    call ModelSpace_elevation_parameters(elev, bounds, genpar)
    wfld%counter = 0

    ! Propagator.f90:168-243
    do it = genpar%tmin, genpar%tmax, genpar%tstep
       if (ndat > 0) then
          select case (ext)
          case (0)
             call Extraction_array_sinc(bounds, model, dat(1:ndat), grid%u2, genpar, it)
          case (1)
             call Extraction_array_simple_afwi_noomp(bounds, model, dat(1:ndat), grid%u2, genpar, it)
          case (2)
             call Extraction_array_LSRTM_sinc(bounds, model, dat(1:ndat), grid%u2, genpar, it)
          case (3)
             call Extraction_array_sinc_afwi(bounds, model, dat(1:ndat), grid%u2, genpar, it)
          end select
       end if
       if (nsnap > 0) call Extraction_wavefield(bounds, model, elev, grid%u2, genpar, it, dat=wfld)
       call Boundary_set_free_surface_grid(bounds, model, elev, grid, genpar)
       if (with_rho /= 0) then
          call FD_2D_derivatives_acoustic_forward_grid(genpar, bounds, grid%u2, grid%u3, model)
       else if (stencil == 0) then
          call FD_2nd_2D_derivatives_scalar_forward_grid(genpar, bounds, grid%u2, grid%u3, model)
       else
          call FD_2nd_2D_derivatives_scalar_adjoint_grid(genpar, bounds, grid%u2, grid%u3, model)
       end if
       if (nsou > 0) then
          select case (inj)
          case (1)
             call Injection_sinc(bounds, model, sou(1:nsou), grid%u3, genpar, it)
          case (2)
             call Injection_simple_noomp(bounds, model, sou(1:nsou), grid%u3, genpar, it)
          case (3)
             call Injection_LSRTM_sinc(bounds, model, sou(1:nsou), grid%u3, genpar, it)
          end select
       end if
       call FD_2nd_time_derivative_grid(genpar, bounds, grid)
       call Boundary_3d(genpar, bounds, grid, model, hig)
       call FDswaptime_pointer(grid)
    end do

    do i = 1, ndat
       dattr(:, i) = dat(i)%trace(:, 1)
    end do
    if (nsnap > 0) then
       do i = 1, nsnap
          wave(:, :, i) = wfld%wave(:, :, 1, i, 1)
       end do
    end if
    call deallocateHigdonParam(hig)
    call deallocateUSpace(grid)
    call deallocateModelSpace_elev(elev)
    if (with_rho /= 0) call deallocate_sxx2d_szz2d_delp2d()
  end subroutine ref_propagate

  ! One 3D forward modeling run in Propagator.f90 order (genpar%twoD false,
  ! nbound 4): ExtractData of kind ext (as ref_propagate), the 3D stencil
  ! FD_2nd_3D_derivatives_scalar_forward_grid, Injection_sinc of the sources,
  ! the time update, Boundary_3d.  vel_pad: the padded model with the bounds
  ! 1-ntaper .. n+ntaper per axis and the 4-cell halos.  with_rho: withRho,
  ! rho_pad the padded density (as vel_pad): Interpolate,
  ! allocate_sxx3d_syy3d_szz3d_delp3d, FD_acoustic_rho_init_coefs with
  ! coefpower 1 and FD_3D_derivatives_acoustic_forward_grid, as
  ! Acoustic_modeling_sep.f90:113-121 calls them.  OpenMP runs on one
  ! thread here: the reference's 3D sinc extraction adds to one trace sample
  ! from every thread of its parallel loop (Extraction.f90:309-319), which is
  ! ordered only on one thread.
  subroutine ref_propagate3d(nz, nxw, nyw, ntaper, dz, dx, dy, dt, oz, oxw, oyw, lsinc, vel_pad, &
       &                     nsou, souz, soux, souy, soutr, ext, ndat, datz, datx, daty, dattr, nt, &
       &                     with_rho, rho_pad, tdir, snapi, nsnap, wave, stencil, inj) bind(C, name='ref_propagate3d')
    use omp_lib
    integer(c_int), value :: nz, nxw, nyw, ntaper, lsinc, nsou, ext, ndat, nt, with_rho, tdir, snapi, nsnap
    integer(c_int), value :: stencil, inj
    real(c_float), value  :: dz, dx, dy, dt, oz, oxw, oyw
    real(c_float)         :: vel_pad(*), souz(*), soux(*), souy(*), soutr(nt, *), datz(*), datx(*), daty(*)
    real(c_float)         :: dattr(nt, *), rho_pad(*), wave(nz, nxw, nyw, *)

    type(GeneralParam)         :: genpar
    type(FDbounds)             :: bounds
    type(ModelSpace)           :: model
    type(ModelSpace_elevation) :: elev
    type(HigdonParam)          :: hig
    type(UnscaledFDcoefs)      :: fdcoefs
    type(ScaledFDcoefs), target :: scaled
    type(USpace)               :: grid
    type(WaveSpace)            :: wfld
    type(TraceSpace), allocatable :: sou(:), dat(:)
    integer :: i, j, k, it, n1, n2, n3, nthr

    genpar%twoD = .false.
    genpar%withRho = with_rho /= 0
    genpar%Born = .false.
    genpar%LSRTM = .false.
    genpar%verbose = .false.
    genpar%optim = .true.
    genpar%dt = dt
    genpar%nt = nt
    genpar%delta = (/ dz, dx, dy /)
    genpar%omodel = (/ oz, oxw, oyw /)
    genpar%lsinc = lsinc
    genpar%nbound = 4
    genpar%ntaper = ntaper
    genpar%snapi = snapi
    genpar%ntsnap = nsnap
    genpar%rec_type = 0
    genpar%shot_type = 0
    genpar%surf_type = 0
    genpar%coefpower = 2
    if (with_rho /= 0) genpar%coefpower = 1
    if (tdir > 0) then
       genpar%tmin = 1; genpar%tmax = nt; genpar%tstep = 1
    else
       genpar%tmin = nt; genpar%tmax = 1; genpar%tstep = -1
    end if

    bounds%nmin1 = -ntaper + 1
    bounds%nmax1 = nz + ntaper
    bounds%nmin2 = -ntaper + 1
    bounds%nmax2 = nxw + ntaper
    bounds%nmin3 = -ntaper + 1
    bounds%nmax3 = nyw + ntaper

    model%nz = nz
    model%nxw = nxw
    model%nyw = nyw
    model%dz = dz
    model%dx = dx
    model%dy = dy
    n1 = bounds%nmax1 - bounds%nmin1 + 9
    n2 = bounds%nmax2 - bounds%nmin2 + 9
    n3 = bounds%nmax3 - bounds%nmin3 + 9
    allocate(model%vel(bounds%nmin1-4:bounds%nmax1+4, bounds%nmin2-4:bounds%nmax2+4, bounds%nmin3-4:bounds%nmax3+4))
    do k = 1, n3
       do j = 1, n2
          do i = 1, n1
             model%vel(bounds%nmin1-5+i, bounds%nmin2-5+j, bounds%nmin3-5+k) = vel_pad(i + (j-1)*n1 + (k-1)*n1*n2)
          end do
       end do
    end do
    if (with_rho /= 0) then
       allocate(model%rho(bounds%nmin1-4:bounds%nmax1+4, bounds%nmin2-4:bounds%nmax2+4, bounds%nmin3-4:bounds%nmax3+4))
       allocate(model%rho2(bounds%nmin1-4:bounds%nmax1+4, bounds%nmin2-4:bounds%nmax2+4, bounds%nmin3-4:bounds%nmax3+4))
       do k = 1, n3
          do j = 1, n2
             do i = 1, n1
                model%rho(bounds%nmin1-5+i, bounds%nmin2-5+j, bounds%nmin3-5+k) = rho_pad(i + (j-1)*n1 + (k-1)*n1*n2)
             end do
          end do
       end do
       call Interpolate(model, bounds, genpar)
    end if

    allocate(sou(max(nsou,1)), dat(max(ndat,1)))
    do i = 1, nsou
       sou(i)%coord = (/ souz(i), soux(i), souy(i) /)
       allocate(sou(i)%trace(nt, 1))
       sou(i)%trace(:, 1) = soutr(:, i)
    end do
    do i = 1, ndat
       dat(i)%coord = (/ datz(i), datx(i), daty(i) /)
       allocate(dat(i)%trace(nt, 1))
       dat(i)%trace(:, 1) = 0.
    end do

    if (nsnap > 0) then
       allocate(wfld%wave(nz, nxw, nyw, nsnap, 1))
       wfld%wave = 0.
    end if
    wfld%counter = 0
    allocate(elev%ielev_z(bounds%nmin2:bounds%nmax2, bounds%nmin3:bounds%nmax3))
    elev%ielev_z = 1

    nthr = omp_get_max_threads()
    call omp_set_num_threads(1)
    call allocateUSpace(grid, genpar, bounds)
    call zeroUSpace(grid)
    allocate(hig%gx(16, bounds%nmin1:bounds%nmax1, bounds%nmin3:bounds%nmax3))
    allocate(hig%gy(16, bounds%nmin1:bounds%nmax1, bounds%nmin2:bounds%nmax2))
    allocate(hig%gz(16, bounds%nmin2:bounds%nmax2, bounds%nmin3:bounds%nmax3))
    if (with_rho /= 0) then
       call FD_acoustic_rho_init_coefs(fdcoefs)
       call allocate_sxx3d_syy3d_szz3d_delp3d(bounds, model)
    else
       call FD_acoustic_init_coefs(fdcoefs)
    end if
    call FD_types_assign_scaled_coefs(fdcoefs, scaled, genpar)
    call FD_derivatives_coef_init(scaled)
    call Higdon(genpar%dt, model, bounds, hig)
    if (nsou > 0) call ModelSpace_compute_array_xyz_positions(genpar, sou(1:nsou))
    if (ndat > 0) call ModelSpace_compute_array_xyz_positions(genpar, dat(1:ndat))

    do it = genpar%tmin, genpar%tmax, genpar%tstep
       if (ndat > 0) then
          select case (ext)
          case (0)
             call Extraction_array_sinc(bounds, model, dat(1:ndat), grid%u2, genpar, it)
          case (1)
             call Extraction_array_simple_afwi_noomp(bounds, model, dat(1:ndat), grid%u2, genpar, it)
          case (2)
             call Extraction_array_LSRTM_sinc(bounds, model, dat(1:ndat), grid%u2, genpar, it)
          case (3)
             call Extraction_array_sinc_afwi(bounds, model, dat(1:ndat), grid%u2, genpar, it)
          end select
       end if
       if (nsnap > 0) call Extraction_wavefield(bounds, model, elev, grid%u2, genpar, it, dat=wfld)
       if (with_rho /= 0) then
          call FD_3D_derivatives_acoustic_forward_grid(genpar, bounds, grid%u2, grid%u3, model)
       else if (stencil == 0) then
          call FD_2nd_3D_derivatives_scalar_forward_grid(genpar, bounds, grid%u2, grid%u3, model)
       else
          call FD_2nd_3D_derivatives_scalar_adjoint_grid(genpar, bounds, grid%u2, grid%u3, model)
       end if
       if (nsou > 0) then
          if (inj == 3) then
             call Injection_LSRTM_sinc(bounds, model, sou(1:nsou), grid%u3, genpar, it)
          else
             call Injection_sinc(bounds, model, sou(1:nsou), grid%u3, genpar, it)
          end if
       end if
       call FD_2nd_time_derivative_grid(genpar, bounds, grid)
       call Boundary_3d(genpar, bounds, grid, model, hig)
       call FDswaptime_pointer(grid)
    end do
    call omp_set_num_threads(nthr)

    do i = 1, ndat
       dattr(:, i) = dat(i)%trace(:, 1)
    end do
    if (nsnap > 0) then
       do i = 1, nsnap
          wave(:, :, :, i) = wfld%wave(:, :, :, i, 1)
       end do
    end if
    call deallocateHigdonParam(hig)
    call deallocateUSpace(grid)
    if (with_rho /= 0) call deallocate_sxx3d_syy3d_szz3d_delp3d()
  end subroutine ref_propagate3d

  ! Interpolate.f90:44-130 on a padded density (as vel_pad): rho2 out
  subroutine ref_interpolate_rho(nz, nxw, ntaper, rho_pad, rho2_pad) bind(C, name='ref_interpolate_rho')
    integer(c_int), value :: nz, nxw, ntaper
    real(c_float)         :: rho_pad(*), rho2_pad(*)
    type(GeneralParam)    :: genpar
    type(FDbounds)        :: bounds
    type(ModelSpace)      :: model
    integer :: i, j, n1, n2
    genpar%nbound = 0
    bounds%nmin1 = -ntaper + 1
    bounds%nmax1 = nz + ntaper
    bounds%nmin2 = -ntaper + 1
    bounds%nmax2 = nxw + ntaper
    bounds%nmin3 = 1
    bounds%nmax3 = 1
    n1 = bounds%nmax1 - bounds%nmin1 + 9
    n2 = bounds%nmax2 - bounds%nmin2 + 9
    allocate(model%rho(bounds%nmin1-4:bounds%nmax1+4, bounds%nmin2-4:bounds%nmax2+4, 1:1))
    allocate(model%rho2(bounds%nmin1-4:bounds%nmax1+4, bounds%nmin2-4:bounds%nmax2+4, 1:1))
    do j = 1, n2
       do i = 1, n1
          model%rho(bounds%nmin1-5+i, bounds%nmin2-5+j, 1) = rho_pad(i + (j-1)*n1)
       end do
    end do
    call Interpolate(model, bounds, genpar)
    do j = 1, n2
       do i = 1, n1
          rho2_pad(i + (j-1)*n1) = model%rho2(bounds%nmin1-5+i, bounds%nmin2-5+j, 1)
       end do
    end do
    deallocate(model%rho, model%rho2)
  end subroutine ref_interpolate_rho

  ! The reference's integer powers of reals as this compiler lowers them
  ! (Imaging.f90:365,470-472: vel**3, rho**2, vel**2): out(:,i) = x**3, x**2,
  ! x**3*y, y**2*x**2
  subroutine ref_powers(n, x, y, out) bind(C, name='ref_powers')
    integer(c_int), value :: n
    real(c_float)         :: x(n), y(n), out(4, n)
    integer :: i
    do i = 1, n
       out(1, i) = x(i)**3
       out(2, i) = x(i)**2
       out(3, i) = x(i)**3*y(i)
       out(4, i) = y(i)**2*x(i)**2
    end do
  end subroutine ref_powers

end module ref_harness
