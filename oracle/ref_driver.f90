! Test harness around the reference's own hydro kernels.
!
! TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile together with the
! reference sources hydro/umuscl.f90, hydro/godunov_utils.f90 and
! hydro/interpol_hydro.f90 (and the modules they use), taken unchanged from
! the reference tree, into oracle/_ref/libramses_ref{1,3}d.so.  Each routine
! below only packs plain C arguments into the reference's own types and calls
! the reference routine, so that tests can compare the C restatement in
! oracle/ramses_oracle.c against the reference arithmetic bit for bit.
module ref_driver
  use iso_c_binding
  use amr_parameters, only: ndim, twotondim, twondim
  use hydro_parameters, only: nvar, nprim, solver_llf, solver_hll, solver_hllc
  use amr_commons, only: hydro_params_t, run_t
  implicit none

contains

  subroutine fill_params(hp, gamma, smallr, smallc, courant_factor, difmag, slope_type, riemann)
    type(hydro_params_t), intent(out) :: hp
    real(c_double), value :: gamma, smallr, smallc, courant_factor, difmag
    integer(c_int), value :: slope_type, riemann
    hp%gamma = gamma
    hp%gamma_rad = 1.33333333334d0
    hp%smallr = smallr
    hp%smallc = smallc
    hp%courant_factor = courant_factor
    hp%difmag = difmag
    hp%slope_type = slope_type
    hp%slope_mag_type = -1
    hp%riemann = riemann
    hp%riemann2d = 0
    hp%switch_llf_dmin = -1
    hp%switch_llf_pmin = -1
  end subroutine fill_params

  ! Number of dimensions this library was compiled for.
  function ref_ndim() bind(C, name="ref_ndim") result(n)
    integer(c_int) :: n
    n = ndim
  end function ref_ndim

  ! unsplit (hydro/umuscl.f90:20) on one stencil block of nn = 2*nsubgrid cells.
  ! uin/gin/flux use the hydro_kernel_t bounds of init_hydro_kernel.
  subroutine ref_unsplit(nn, uin, gin, flux, dx, dt, gamma, smallr, smallc, courant_factor, &
       & difmag, slope_type, riemann) bind(C, name="ref_unsplit")
    integer(c_int), value :: nn, slope_type, riemann
    real(c_double), value :: dx, dt, gamma, smallr, smallc, courant_factor, difmag
    real(c_double), dimension(*) :: uin, gin, flux
    type(hydro_params_t) :: hp
    integer :: iu1, iu2, ju1, ju2, ku1, ku2, if1, if2, jf1, jf2, kf1, kf2
    integer :: nu, nf
    real(kind=8), allocatable :: qin(:), cin(:), dq(:), qm(:), qp(:), fx(:), tx(:), divu(:), tmp(:)

    call fill_params(hp, gamma, smallr, smallc, courant_factor, difmag, slope_type, riemann)
    iu1 = -1; iu2 = nn + 2
    ju1 = (1 - ndim/2) - 1*(ndim/2); ju2 = (1 - ndim/2) + (nn + 2)*(ndim/2)
    ku1 = (1 - ndim/3) - 1*(ndim/3); ku2 = (1 - ndim/3) + (nn + 2)*(ndim/3)
    if1 = 1; if2 = nn + 1
    jf1 = 1; jf2 = (1 - ndim/2) + (nn + 1)*(ndim/2)
    kf1 = 1; kf2 = (1 - ndim/3) + (nn + 1)*(ndim/3)
    nu = (iu2 - iu1 + 1)*(ju2 - ju1 + 1)*(ku2 - ku1 + 1)
    nf = (if2 - if1 + 1)*(jf2 - jf1 + 1)*(kf2 - kf1 + 1)
    allocate(qin(nu*nprim), cin(nu), dq(nu*nprim*ndim), qm(nu*nprim*ndim), qp(nu*nprim*ndim))
    allocate(fx(nu*nprim), tx(nu*2), divu(nf), tmp(nf*2*ndim))
    qin = 0; cin = 0; dq = 0; qm = 0; qp = 0; fx = 0; tx = 0; divu = 0; tmp = 0
    call unsplit(uin, gin, qin, cin, flux, tmp, dq, qm, qp, fx, tx, divu, dx, dt, &
         & iu1, iu2, ju1, ju2, ku1, ku2, if1, if2, jf1, jf2, kf1, kf2, hp)
    deallocate(qin, cin, dq, qm, qp, fx, tx, divu, tmp)
  end subroutine ref_unsplit

  ! riemann_llf / riemann_hll / riemann_hllc (hydro/godunov_utils.f90:338-857)
  subroutine ref_riemann(solver, ql, qr, f, gamma, smallr, smallc, smax_cap) bind(C, name="ref_riemann")
    integer(c_int), value :: solver
    real(c_double), value :: gamma, smallr, smallc, smax_cap
    real(c_double), dimension(nprim) :: ql, qr
    real(c_double), dimension(nprim + 1) :: f
    type(hydro_params_t) :: hp
    call fill_params(hp, gamma, smallr, smallc, 0.5d0, 0.0d0, 1, solver)
    hp%smax_cap = smax_cap
    if (solver == solver_llf) then
       call riemann_llf(ql, qr, f, hp)
    else if (solver == solver_hll) then
       call riemann_hll(ql, qr, f, hp)
    else
       call riemann_hllc(ql, qr, f, hp)
    end if
  end subroutine ref_riemann

  ! cmpdt (hydro/godunov_utils.f90:228); uu is modified as by the reference
  subroutine ref_cmpdt(uu, gg, dx, gamma, smallr, smallc, courant_factor, dt) bind(C, name="ref_cmpdt")
    real(c_double), dimension(nvar) :: uu
    real(c_double), dimension(ndim) :: gg
    real(c_double), value :: dx, gamma, smallr, smallc, courant_factor
    real(c_double) :: dt
    type(run_t) :: r
    real(kind=8), dimension(1:6) :: bb
    r%gamma = gamma
    r%smallr = smallr
    r%smallc = smallc
    r%courant_factor = courant_factor
    bb = 0.0d0
    call cmpdt(r, uu, bb, gg, dx, dt)
  end subroutine ref_cmpdt

  ! interpol_hydro (hydro/interpol_hydro.f90:5)
  subroutine ref_interpol_hydro(u1, u2, interpol_var, interpol_type, smallr) bind(C, name="ref_interpol_hydro")
    real(c_double), dimension(0:twondim, 1:nvar) :: u1
    real(c_double), dimension(1:twotondim, 1:nvar) :: u2
    integer(c_int), value :: interpol_var, interpol_type
    real(c_double), value :: smallr
    integer :: iv, it
    real(kind=8) :: sr
    iv = interpol_var; it = interpol_type; sr = smallr
    call interpol_hydro(u1, u2, iv, it, sr)
  end subroutine ref_interpol_hydro

end module ref_driver
