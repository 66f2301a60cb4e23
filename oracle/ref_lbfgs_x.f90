! ref_lbfgs_x.f90 — test infrastructure, not part of the product.
!
! One run of progs/inversion/L_BFGS.f90 with lbfgs_type 1, without SEPlib:
! the program's own file handling (nlinv_read_array_params.f90:139-174 names
! and control files, :237-248 DIAG and W reads; L_BFGS.f90:69-90 writes) around
! the reference's LBFGSOC (libs/libnlinv/nlinv_lbfgs.f, linked beside this
! file in oracle/_ref/libnlinvref.so), with x, f and g passed in place of the
! SEP files mod, fct and gdt.  The files are opened and written by this
! compiler's Fortran runtime, so tests can pin the file forms a C reader and
! writer must match.  The iteration log (IPRINT(1) >= 0: LB6 on unit MP) is
! switched off; it writes nothing the next run reads.
module ref_lbfgs_x_mod
  use iso_c_binding
  implicit none
contains

  subroutine ref_lbfgs_x(cdir, ldir, n, m, x, f, g, nparam, xmin, xmax, stp1_opt, stp_init, boundt, &
       &                 iflag_out, info_out) bind(C, name='ref_lbfgs_x')
    integer(c_int), value :: ldir, n, m, nparam, stp1_opt, boundt
    character(kind=c_char), intent(in) :: cdir(ldir)
    real(c_double), intent(inout) :: x(n)
    real(c_double), value :: f, stp_init
    real(c_double), intent(in) :: g(n), xmin(nparam), xmax(nparam)
    integer(c_int), intent(out) :: iflag_out, info_out

    character(len=1024) :: dir, lbin, lbout, mcin, mcout
    double precision, allocatable :: w(:), diag(:), gd(:)
    double precision :: fd, eps, xtol, stpi
    integer :: iflag, myinfo, iprint(2), nwork, k
    logical :: ex
    external :: lbfgsoc

    dir = ''
    do k = 1, ldir
       dir(k:k) = cdir(k)
    end do
    lbin = trim(dir)//'/LbfgsDat.dat'
    lbout = trim(dir)//'/LbfgsDatOut.dat'
    mcin = trim(dir)//'/McsrchDat.dat'
    mcout = trim(dir)//'/McsrchDatOut.dat'

    inquire(file=trim(dir)//'/Iflag.dat', exist=ex)
    iflag = 0
    if (ex) then
       open(10, file=trim(dir)//'/Iflag.dat')
       read(10, *) iflag
       close(10)
    end if
    inquire(file=trim(dir)//'/Info.dat', exist=ex)
    myinfo = 0
    if (ex) then
       open(40, file=trim(dir)//'/Info.dat')
       read(40, *) myinfo
       close(40)
    end if

    nwork = n*(2*m+1) + 2*m
    allocate(w(nwork), diag(n), gd(n))
    w = 0.d0
    diag = 0.d0
    if (iflag .ne. 0) then
       open(88, file=trim(dir)//'/Diag.dat', form='UNFORMATTED')
       rewind(88)
       read(88) diag
       close(88)
       open(89, file=trim(dir)//'/WorkArray.dat', form='UNFORMATTED')
       rewind(89)
       read(89) w
       close(89)
    end if

    gd = g
    fd = f
    eps = 1.0d-20
    xtol = epsilon(xtol)
    stpi = stp_init
    iprint(1) = -1
    iprint(2) = 0
    call lbfgsoc(n, m, x, fd, gd, .false., diag, iprint, eps, xtol, stpi, w, iflag, myinfo, &
         &       lbin, lbout, mcin, mcout, xmin, xmax, nparam, stp1_opt, boundt)

    open(89, file=trim(dir)//'/WorkArrayOut.dat', status='replace', form='UNFORMATTED')
    rewind(89)
    write(89) w
    close(89)
    open(88, file=trim(dir)//'/DiagOut.dat', status='replace', form='UNFORMATTED')
    rewind(88)
    write(88) diag
    close(88)
    open(30, file=trim(dir)//'/IflagOut.dat', status='replace')
    rewind(30)
    write(30, "(I2)") iflag
    close(30)
    open(20, file=trim(dir)//'/InfoOut.dat', status='replace')
    rewind(20)
    write(20, "(I2)") myinfo
    close(20)

    iflag_out = iflag
    info_out = myinfo
    deallocate(w, diag, gd)
  end subroutine ref_lbfgs_x

end module ref_lbfgs_x_mod
