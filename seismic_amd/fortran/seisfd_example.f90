! seisfd_example.f90 -- a Fortran host program driving the engine through
! seisfd_mod, the way the reference's drivers would: forward modeling of one
! shot (Acoustic_modeling_sep.f90:80-88) and an l_bfgs run with a Fortran
! fctgdt callback (Inversion.f90:29-189).
!
! usage: seisfd_example OUTDIR
!   writes OUTDIR/vel.f32 (nz,nx), OUTDIR/wavelet.f32 (nt), OUTDIR/gather.f32
!   (nt,ntr) and OUTDIR/lbfgs.txt (iterations, evaluations, final misfit).
!
! usage: seisfd_example --lbfgsoc DIR
!   one model update of the out-of-core L-BFGS per process, as
!   progs/inversion/L_BFGS.f90 runs LBFGSOC: reads DIR/opts.txt (n m nparam
!   stp1_opt step_init clip_type, then xmin(1:nparam), xmax(1:nparam)),
!   DIR/ctl.txt (iflag info; absent = 0 0, like Iflag.dat / Info.dat),
!   DIR/model.f64 (n doubles) and DIR/fg.f64 (f, then n doubles of g);
!   keeps the optimiser state in DIR/state.bin and writes the next model to
!   DIR/model.f64 and "iflag info" to DIR/ctl.txt.
module example_fct
  use iso_c_binding
  implicit none
  integer, parameter :: nq = 16
  real(c_double) :: xstar(nq)
contains
  ! f = 1/2 sum k (x_k - x*_k)^2 ; returns g = -df/dx (the reference's sign)
  integer(c_int) function quad_fct(ctx, x, g, f) bind(C)
    type(c_ptr), value :: ctx
    real(c_float), intent(in) :: x(nq)
    real(c_float), intent(out) :: g(nq)
    real(c_double), intent(out) :: f
    integer :: k
    f = 0d0
    do k = 1, nq
       f = f + 0.5d0 * k * (dble(x(k)) - xstar(k))**2
       g(k) = real(-k * (dble(x(k)) - xstar(k)), c_float)
    end do
    quad_fct = 0
  end function quad_fct
end module example_fct

program seisfd_example
  use iso_c_binding
  use seisfd_mod
  use example_fct
  implicit none
  integer, parameter :: nz = 60, nx = 80, nt = 300, ntr = 40
  real(c_float), target :: vel(nz, nx), wavelet(nt), rz(ntr), rx(ntr), gather(nt, ntr)
  real(c_float) :: x(nq)
  type(sfd_grid) :: g
  type(sfd_shot) :: shot(1)
  type(sfd_lbfgs_opts) :: lo
  type(sfd_lbfgs_report) :: rep
  type(sfd_fwi_opts) :: fo
  type(sfd_lbfgsoc_opts) :: oco
  type(sfd_grid3d) :: g3
  type(sfd_shot3d) :: s3
  type(c_ptr) :: h
  character(len=512) :: outdir
  integer :: i, j, rc, u
  real :: t, a, pi

  call get_command_argument(1, outdir)
  if (trim(outdir) == '--sizes') then
     ! struct sizes as Fortran sees them (checked against C by the tests)
     write(*, '(8i6)') c_sizeof(g), c_sizeof(shot(1)), c_sizeof(lo), c_sizeof(rep), c_sizeof(fo), &
          c_sizeof(oco), c_sizeof(g3), c_sizeof(s3)
     stop
  end if
  if (trim(outdir) == '--lbfgsoc') then
     call get_command_argument(2, outdir)
     call lbfgsoc_update(trim(outdir))
     stop
  end if
  pi = 3.14159265
  ! three flat layers
  do j = 1, nx
     do i = 1, nz
        vel(i, j) = 1500.0
        if (i > 20) vel(i, j) = 2000.0
        if (i > 40) vel(i, j) = 2500.0
     end do
  end do
  ! 15 Hz Ricker, delay 1.2/f
  do i = 1, nt
     t = (i - 1) * 1e-3 - 1.2 / 15.0
     a = (pi * 15.0 * t)**2
     wavelet(i) = (1.0 - 2.0 * a) * exp(-a)
  end do
  do i = 1, ntr
     rz(i) = 20.0
     rx(i) = 20.0 * (i - 1)
  end do

  g%nz = nz; g%nx = nx; g%oz = 0.0; g%ox = 0.0; g%dz = 10.0; g%dx = 10.0; g%dt = 1e-3
  g%ntaper = 20; g%lsinc = 7; g%snapi = 4; g%aperture_x = 0.0
  g%rec_type = 0; g%shot_type = 0; g%surf_type = 0
  rc = sfd_create(h, g)
  if (rc /= SFD_OK) stop 'sfd_create failed'
  rc = sfd_set_model(h, vel, c_null_ptr)
  if (rc /= SFD_OK) then
     write(0, *) trim(sfd_error_message(h))
     stop 'sfd_set_model failed'
  end if
  shot(1)%sz = 20.0; shot(1)%sx = 400.0; shot(1)%nt = nt; shot(1)%ntr = ntr
  shot(1)%wavelet = c_loc(wavelet); shot(1)%rz = c_loc(rz); shot(1)%rx = c_loc(rx)
  gather = 0.0
  rc = sfd_forward(h, shot, 1, gather, SFD_EXTRACT_SINC_V2)
  if (rc /= SFD_OK) then
     write(0, *) trim(sfd_error_message(h))
     stop 'sfd_forward failed'
  end if
  rc = sfd_destroy(h)

  open(newunit=u, file=trim(outdir)//'/vel.f32', access='stream', form='unformatted', status='replace')
  write(u) vel
  close(u)
  open(newunit=u, file=trim(outdir)//'/wavelet.f32', access='stream', form='unformatted', status='replace')
  write(u) wavelet
  close(u)
  open(newunit=u, file=trim(outdir)//'/gather.f32', access='stream', form='unformatted', status='replace')
  write(u) gather
  close(u)

  ! l_bfgs with a Fortran callback
  do i = 1, nq
     xstar(i) = 2000d0 + 50d0 * i
     x(i) = 2000.0
  end do
  lo%niter = 20; lo%neval = 60; lo%m = 5; lo%const_type = 2; lo%freeze_soft = 1; lo%eval0 = 0
  lo%xmin = 1500d0; lo%xmax = 4500d0
  rep%hist_cap = 0
  rep%f_eval = c_null_ptr; rep%t_eval = c_null_ptr; rep%accepted = c_null_ptr
  rc = sfd_lbfgs_run(lo, nq, x, c_null_ptr, c_null_ptr, c_funloc(quad_fct), c_null_ptr, rep)
  if (rc /= SFD_OK) stop 'sfd_lbfgs_run failed'
  open(newunit=u, file=trim(outdir)//'/lbfgs.txt', status='replace')
  write(u, '(3i8, es24.15)') rep%iterations, rep%evaluations, rep%iflag, maxval(abs(dble(x) - xstar))
  close(u)
  write(*, '(a, i0, a, i0, a, es10.3)') 'seisfd_example: gather done; l_bfgs iterations ', rep%iterations, &
       ' evaluations ', rep%evaluations, ' max|x-x*| ', maxval(abs(dble(x) - xstar))
contains

  ! One LBFGSOC update (progs/inversion/L_BFGS.f90:54-98) on the files of dir.
  subroutine lbfgsoc_update(dir)
    character(len=*), intent(in) :: dir
    type(sfd_lbfgsoc_opts) :: o
    real(c_double), allocatable, target :: xm(:), gd(:), bmin(:), bmax(:)
    real(c_double) :: fv, step_init
    integer(c_int) :: n, m, nparam, stp1_opt, clip_type, iflag, info, rc
    character(len=1024) :: state
    logical :: have
    integer :: u
    open(newunit=u, file=dir//'/opts.txt', status='old')
    read(u, *) n, m, nparam, stp1_opt, step_init, clip_type
    allocate(xm(n), gd(n), bmin(nparam), bmax(nparam))
    read(u, *) bmin
    read(u, *) bmax
    close(u)
    iflag = 0
    info = 0
    inquire(file=dir//'/ctl.txt', exist=have)
    if (have) then
       open(newunit=u, file=dir//'/ctl.txt', status='old')
       read(u, *) iflag, info
       close(u)
    end if
    open(newunit=u, file=dir//'/model.f64', access='stream', form='unformatted', status='old')
    read(u) xm
    close(u)
    open(newunit=u, file=dir//'/fg.f64', access='stream', form='unformatted', status='old')
    read(u) fv, gd
    close(u)
    o%m = m
    o%nparam = nparam
    o%xmin = c_loc(bmin)
    o%xmax = c_loc(bmax)
    o%stp1_opt = stp1_opt
    o%step_init = step_init
    o%clip_type = clip_type
    state = dir//'/state.bin'//c_null_char
    rc = sfd_lbfgsoc_step(state, state, o, n, xm, fv, gd, iflag, info)
    if (rc /= 0) then
       write(0, *) 'sfd_lbfgsoc_step failed: ', rc
       error stop 2
    end if
    open(newunit=u, file=dir//'/model.f64', access='stream', form='unformatted', status='replace')
    write(u) xm
    close(u)
    open(newunit=u, file=dir//'/ctl.txt', status='replace')
    write(u, '(2i4)') iflag, info
    close(u)
  end subroutine lbfgsoc_update

end program seisfd_example
