! ref_harness.f90 -- C-callable entry points into reference module routines
! that the reference itself does not export with BIND(C).
!
! TEST INFRASTRUCTURE ONLY.  Linked into oracle/_ref/libfteikref.so together
! with the reference Fortran compiled from /root/reference (see Makefile), and
! used only by tests/golden/gen_golden.py and the oracle-pinning tests.
! Each wrapper forwards its arguments unchanged to the reference routine it
! names; no arithmetic happens here.

! localSolver.F90:878-959 (needs model + source module state set through the
! reference's own BIND(C) setters first).
subroutine refh_ls3d_initialize(isrc, eps, dest, ts8, ierr) &
           bind(c, name='refh_ls3d_initialize')
  use iso_c_binding
  use fteik_localsolver3d64f, only : fteik_localSolver_initialize64fF
  implicit none
  integer(c_int), value :: isrc
  real(c_double), value :: eps
  integer(c_int), intent(out) :: dest(8), ierr
  real(c_double), intent(out) :: ts8(8)
  integer d(8), ie
  call fteik_localSolver_initialize64fF(isrc, eps, d, ts8, ie)
  dest(:) = d(:)
  ierr = ie
end subroutine

! localSolver.F90:384-534
subroutine refh_ls3d_noinit(n, tv, te, tn, tev, ten, tnv, tnve, &
                            s1, s2, s3, s4, s5, s6, s7, tupd) &
           bind(c, name='refh_ls3d_noinit')
  use iso_c_binding
  use fteik_localsolver3d64f, only : fteik_localSolver_noInit64fF
  implicit none
  integer(c_int), value :: n
  real(c_double), intent(in) :: tv(n), te(n), tn(n), tev(n), ten(n), tnv(n), tnve(n)
  real(c_double), intent(in) :: s1(n), s2(n), s3(n), s4(n), s5(n), s6(n), s7(n)
  real(c_double), intent(inout) :: tupd(n)
  call fteik_localSolver_noInit64fF(n, tv, te, tn, tev, ten, tnv, tnve, &
                                    s1, s2, s3, s4, s5, s6, s7, tupd)
end subroutine

! localSolver.F90:538-751
function refh_ls3d_init(tv, te, tn, tev, ten, tnv, tnve, tt0, &
                        s1, s2, s3, s4, s5, s6, s7, i, j, k, &
                        sgntz, sgntx, sgnty, sgnrz_dzi, sgnrx_dxi, sgnry_dyi) &
         result(tupd) bind(c, name='refh_ls3d_init')
  use iso_c_binding
  use fteik_localsolver3d64f, only : fteik_localSolver_init64fF
  implicit none
  real(c_double), value :: tv, te, tn, tev, ten, tnv, tnve, tt0
  real(c_double), value :: s1, s2, s3, s4, s5, s6, s7
  integer(c_int), value :: i, j, k, sgntz, sgntx, sgnty
  real(c_double), value :: sgnrz_dzi, sgnrx_dxi, sgnry_dyi
  real(c_double) :: tupd
  tupd = fteik_localSolver_init64fF(tv, te, tn, tev, ten, tnv, tnve, tt0, &
                                    s1, s2, s3, s4, s5, s6, s7, i, j, k,  &
                                    sgntz, sgntx, sgnty,                  &
                                    sgnrz_dzi, sgnrx_dxi, sgnry_dyi)
end function

! localSolver.F90:299-352
subroutine refh_ls2d_initialize(isrc, eps, dest, ts4, ierr) &
           bind(c, name='refh_ls2d_initialize')
  use iso_c_binding
  use fteik_localsolver2d64f, only : fteik_localSolver2d_initialize64f
  implicit none
  integer(c_int), value :: isrc
  real(c_double), value :: eps
  integer(c_int), intent(out) :: dest(4), ierr
  real(c_double), intent(out) :: ts4(4)
  integer d(4), ie
  call fteik_localSolver2d_initialize64f(isrc, eps, d, ts4, ie)
  dest(:) = d(:)
  ierr = ie
end subroutine

! localSolver.F90:123-158 (ttvec, sloc packed by 4 per node)
subroutine refh_ls2d_noinit(n, ttvec, sloc, tupd) bind(c, name='refh_ls2d_noinit')
  use iso_c_binding
  use fteik_localsolver2d64f, only : fteik_localSolver2d_noInit64f
  implicit none
  integer(c_int), value :: n
  real(c_double), intent(in) :: ttvec(4*n), sloc(4*n)
  real(c_double), intent(out) :: tupd(n)
  call fteik_localSolver2d_noInit64f(n, ttvec, sloc, tupd)
end subroutine

! localSolver.F90:195-289
subroutine refh_ls2d_init(sweep, nz, nx, iz, ix, slow, tt, tupd) &
           bind(c, name='refh_ls2d_init')
  use iso_c_binding
  use fteik_localsolver2d64f, only : fteik_localSolver2d_init64f, &
                                     fteik_localSolver2d_getSweepSigns
  implicit none
  integer(c_int), value :: sweep, nz, nx, iz, ix
  real(c_double), intent(in) :: slow((nz-1)*(nx-1)), tt(nz*nx)
  real(c_double), intent(out) :: tupd
  integer sgntz, sgntx, sgnvz, sgnvx, ierr
  call fteik_localSolver2d_getSweepSigns(sweep, sgntz, sgntx, sgnvz, sgnvx, ierr)
  call fteik_localSolver2d_init64f(sgntz, sgntx, sgnvz, sgnvx, nz, nx, iz, ix, &
                                   slow, tt, tupd)
end subroutine

! refsrc/fteik2d.F90 (original Noble-Gesret 2D solver)
subroutine refh_orig_solver2d(slow, tt, nz, nx, zsrc, xsrc, dz, dx, nsweep) &
           bind(c, name='refh_orig_solver2d')
  use iso_c_binding
  use fteik2d, only : solver2d
  implicit none
  integer(c_int), value :: nz, nx, nsweep
  real(c_double), value :: zsrc, xsrc, dz, dx
  real(c_double), intent(in) :: slow(nz-1, nx-1)
  real(c_double), intent(out) :: tt(nz, nx)
  call solver2d(slow, tt, nz, nx, zsrc, xsrc, dz, dx, nsweep)
end subroutine

! refsrc/fteik.f90 (original Noble-Gesret 3D solver, velocity in cells)
subroutine refh_orig_solver3d(vel, tt, nz, nx, ny, zs, xs, ys, dz, dx, dy, &
                              nsweep, eps) bind(c, name='refh_orig_solver3d')
  use iso_c_binding
  implicit none
  integer(c_int), value :: nz, nx, ny, nsweep
  real(c_double), value :: zs, xs, ys, dz, dx, dy, eps
  real(c_double), intent(in) :: vel(nz-1, nx-1, ny-1)
  real(c_double), intent(out) :: tt(nz, nx, ny)
  interface
    subroutine FTeik(vel, tt, nz, nx, ny, zsin, xsin, ysin, dzin, dxin, dyin, &
                     nsweep, epsin)
      integer, intent(in) :: nz, nx, ny, nsweep
      real*8, intent(in) :: dzin, dxin, dyin, zsin, xsin, ysin, epsin
      real*8, intent(out) :: tt(nz, nx, ny)
      real*8, intent(in) :: vel(nz-1, nx-1, ny-1)
    end subroutine
  end interface
  call FTeik(vel, tt, nz, nx, ny, zs, xs, ys, dz, dx, dy, nsweep, eps)
end subroutine

! solver2D.F90:984-1032 through the reference's own call path:
! fteik_solver2d_traceRaysToPoints (solver2D.F90:310-355) differentiates the
! solved field before it checks the ray destinations, so a call with one
! destination outside the model computes tgradx/tgradz and returns with
! ierr = 1 before any tracing.  The gradients are copied out unchanged.
subroutine refh_solver2d_gradient(n, gx, gz, ierr) &
           bind(c, name='refh_solver2d_gradient')
  use iso_c_binding
  use fteik2d_solver64f, only : fteik_solver2d_traceRaysToPoints, &
                                tgradx, tgradz
  implicit none
  integer(c_int), value :: n
  real(c_double), intent(out) :: gx(n), gz(n)
  integer(c_int), intent(out) :: ierr
  real(c_double) :: xp(1), zp(1)
  xp(1) = -1.d30
  zp(1) = -1.d30
  call fteik_solver2d_traceRaysToPoints(1, xp, zp, ierr)
  gx(1:n) = tgradx(1:n)
  gz(1:n) = tgradz(1:n)
end subroutine
