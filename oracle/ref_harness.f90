! ref_harness.f90 -- test infrastructure for the oracle, not part of the product.
!
! Runs FFT-free pieces of the reference solver, compiled from the reference's
! own sources (oracle/Makefile), on inputs written by tests/test_oracle_ref.py,
! so that the numpy oracle can be compared with the reference's arithmetic:
!
!   ref_harness ftvspline  <dir>   FT_V_spline (FT_V_spline.f90:14-276) through
!                                  Select_Pot (V_impur.f90:84-929)
!   ref_harness fforma     <dir>   fvlj4 (fforma.f90:11-61, S13ADF)
!   ref_harness initcg     <dir>   wcgk (initcg.f90:5-51)
!   ref_harness deriv      <dir>   derivnD with npd=13, icon=13
!                                  (init_deriv_parallel.f90, redef.f, newder.f)
!   ref_harness ironing    <dir>   ironing (ironing.f90:4-33)
!   ref_harness printoutc  <dir>   printoutc (printoutc.f90:13-124) writing den.out/imp.out
!   ref_harness readenc    <dir>   readenc mode 2 (readenc.f90:96-114) reading den.in/imp.in
!                                  (run with <dir> as the working directory: the
!                                  file names are character(len=60))
!
! <dir>/params.txt holds the scalar inputs (list-directed); arrays travel as
! raw little-endian stream files in Fortran order.
program ref_harness
  use Para_DerivnD
  use grid, only: nx, ny, nz
  use gridk, only: pmod
  use util1, only: pi, twopi
  implicit none
  character(len=32) :: cmd
  character(len=512) :: dir
  real(kind=8), external :: FT_V_spline
  integer :: u

  call get_command_argument(1, cmd)
  call get_command_argument(2, dir)
  pi = 4.0d0 * datan(1.0d0)
  twopi = 2.0d0 * pi
  open(newunit=u, file=trim(dir) // '/params.txt', status='old', action='read')

  select case (trim(cmd))
  case ('ftvspline')
    call run_ftvspline()
  case ('fforma')
    call run_fforma()
  case ('initcg')
    call run_initcg()
  case ('deriv')
    call run_deriv()
  case ('ironing')
    call run_ironing()
  case ('printoutc')
    call run_printoutc()
  case ('readenc')
    call run_readenc()
  case default
    write(0, *) 'unknown command ', trim(cmd)
    error stop 2
  end select
  close(u)

contains

  subroutine read_array(name, a, n)
    character(len=*), intent(in) :: name
    integer, intent(in) :: n
    real(kind=8), intent(out) :: a(n)
    integer :: v
    open(newunit=v, file=trim(dir) // '/' // name, access='stream', form='unformatted', status='old')
    read(v) a
    close(v)
  end subroutine read_array

  subroutine write_array(name, a, n)
    character(len=*), intent(in) :: name
    integer, intent(in) :: n
    real(kind=8), intent(in) :: a(n)
    integer :: v
    open(newunit=v, file=trim(dir) // '/' // name, access='stream', form='unformatted', status='replace')
    write(v) a
    close(v)
  end subroutine write_array

  ! params: selec  r_cutoff  umax  pmax  hp  np ; p.bin -> vq.bin
  subroutine run_ftvspline()
    character(len=80) :: selec
    real(kind=8) :: r_cutoff, umax, pmax, hp
    integer :: np, i
    real(kind=8), allocatable :: p(:), v(:)
    read(u, *) selec, r_cutoff, umax, pmax, hp, np
    allocate(p(np), v(np))
    call read_array('p.bin', p, np)
    do i = 1, np
      v(i) = FT_V_spline(p(i), pmax, hp, selec, r_cutoff, umax)
    end do
    call write_array('vq.bin', v, np)
  end subroutine run_ftvspline

  ! params: nx ny nz eps sigma h ; pmod.bin -> fvlj4.bin
  subroutine run_fforma()
    real(kind=8) :: eps, sigma, h, b
    real(kind=8), allocatable :: pm(:), fv(:)
    integer :: m
    read(u, *) nx, ny, nz, eps, sigma, h
    m = (nx / 2 + 1) * ny * nz
    allocate(pm(m), fv(m))
    call read_array('pmod.bin', pm, m)
    b = 0.0d0
    call fforma('OT', b, eps, sigma, h, nx, ny, nz, pm, fv)
    call write_array('fvlj4.bin', fv, m)
  end subroutine run_fforma

  ! params: nx ny nz h ; pmod.bin -> wcgk.bin
  subroutine run_initcg()
    real(kind=8) :: h
    real(kind=8), allocatable :: wk(:)
    integer :: m
    read(u, *) nx, ny, nz, h
    m = (nx / 2 + 1) * ny * nz
    allocate(pmod(nx / 2 + 1, ny, nz), wk(m))
    call read_array('pmod.bin', pmod, m)
    call initcg(h, wk)
    call write_array('wcgk.bin', wk, m)
  end subroutine run_initcg

  ! params: nx ny nz id iv h is_complex ; f.bin -> df.bin (real or complex)
  subroutine run_deriv()
    integer :: id, iv, iscx, nn(3), m
    real(kind=8) :: h
    real(kind=8), allocatable :: f(:,:,:), df(:,:,:), raw(:)
    complex(kind=8), allocatable :: fc(:,:,:), dfc(:,:,:)
    read(u, *) nn(1), nn(2), nn(3), id, iv, h, iscx
    m = nn(1) * nn(2) * nn(3)
    call Init_deriv_p(13, 2, 1)
    if (iscx == 0) then
      allocate(f(nn(1), nn(2), nn(3)), df(nn(1), nn(2), nn(3)))
      call read_array('f.bin', f, m)
      call derivnD(id, nn, h, iv, f, df, 13)
      call write_array('df.bin', df, m)
    else
      allocate(fc(nn(1), nn(2), nn(3)), dfc(nn(1), nn(2), nn(3)), raw(2 * m))
      call read_array('f.bin', raw, 2 * m)
      fc = reshape(cmplx(raw(1::2), raw(2::2), kind=8), nn)
      call derivnD(id, nn, h, iv, fc, dfc, 13)
      raw(1::2) = reshape(real(dfc, kind=8), [m])
      raw(2::2) = reshape(aimag(dfc), [m])
      call write_array('df.bin', raw, 2 * m)
    end if
  end subroutine run_deriv

  ! params: nx ny nz passes ; f.bin -> out.bin
  subroutine run_ironing()
    integer :: n1, n2, n3, passes, k
    real(kind=8), allocatable :: f(:,:,:)
    read(u, *) n1, n2, n3, passes
    allocate(f(n1, n2, n3))
    call read_array('f.bin', f, n1 * n2 * n3)
    do k = 1, passes
      call ironing(f, n1, n2, n3)
    end do
    call write_array('out.bin', f, n1 * n2 * n3)
  end subroutine run_ironing

  ! params: nx ny nz n_par xmax ymax zmax hx hy hz time iprint iter paflv ;
  ! psi.bin rimp.bin vimp.bin -> den.out imp.out
  subroutine run_printoutc()
    integer :: n1, n2, n3, np, iprint, iter, m
    real(kind=8) :: xm, ym, zm, h1, h2, h3, time, paflv
    real(kind=8), allocatable :: raw(:), ri(:), vi(:)
    complex(kind=8), allocatable :: ps(:,:,:)
    character(len=60) :: f1, f2
    read(u, *) n1, n2, n3, np, xm, ym, zm, h1, h2, h3, time, iprint, iter, paflv
    m = n1 * n2 * n3
    allocate(raw(2 * m), ps(n1, n2, n3), ri(3 * np), vi(3 * np))
    call read_array('psi.bin', raw, 2 * m)
    ps = reshape(cmplx(raw(1::2), raw(2::2), kind=8), [n1, n2, n3])
    call read_array('rimp.bin', ri, 3 * np)
    call read_array('vimp.bin', vi, 3 * np)
    f1 = 'den.out'
    f2 = 'imp.out'
    call printoutc(time, iprint, f1, f2, ps, n1, n2, n3, h1, h2, h3, xm, ym, zm, &
                   reshape(ri, [np, 3]), reshape(vi, [np, 3]), np, paflv, iter)
  end subroutine run_printoutc

  ! params: nx ny nz n_par ; den.in imp.in -> psi.bin rimp.bin vimp.bin
  subroutine run_readenc()
    use rho, only: psi, den
    use impur, only: N_par, rimp, vimp
    integer :: n1, n2, n3, m
    real(kind=8) :: r_clust
    real(kind=8), allocatable :: raw(:)
    character(len=60) :: f1, f2
    read(u, *) n1, n2, n3, N_par
    nx = n1
    ny = n2
    nz = n3
    m = n1 * n2 * n3
    allocate(psi(n1, n2, n3), den(n1, n2, n3), rimp(N_par, 3), vimp(N_par, 3), raw(2 * m))
    f1 = 'den.in'
    f2 = 'imp.in'
    call readenc(100, 0.0218d0, f1, f2, 2, 0.d0, r_clust)
    raw(1::2) = reshape(real(psi, kind=8), [m])
    raw(2::2) = reshape(aimag(psi), [m])
    call write_array('psi.bin', raw, 2 * m)
    call write_array('rimp.bin', reshape(rimp, [3 * N_par]), 3 * N_par)
    call write_array('vimp.bin', reshape(vimp, [3 * N_par]), 3 * N_par)
  end subroutine run_readenc

end program ref_harness
