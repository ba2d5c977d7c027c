! Sod shock tube driven from Fortran through hip_interface, in the order of
! calls a RAMSES level step makes (amr/amr_step.f90:276-455): newdt_fine ->
! courant_fine (rh_cmpdt), set_unew, godunov_fine (rh_godunov), set_uold; the
! set-up mirrors metal_allocate_amr / r_set_grid_device / metal_set_nbor
! (metal/metal_runner.f90:13-134,520-560).
!
! 1024 cells on [0,1] (level 10, 512 octs), gamma=1.4, MinMod, HLLC,
! Courant 0.8, zero-gradient ends, t_end=0.2 landed as pm/newdt_fine.f90:156-180
! with nlanding=3.  Writes nstep, t and uold(2,5,512) to the file named by the
! first argument (stream, native endianness).
program sod_driver
  use iso_c_binding
  use hip_interface
  implicit none

  integer, parameter :: nlevel = 10, noct = 512, nvar = 5, ttd = 2
  type, bind(C) :: oct1_t
     integer(c_int64_t) :: hkey
     integer(c_int) :: ckey(1)
     integer(c_int8_t) :: refined(2)
     integer(c_int) :: lev
  end type oct1_t

  type(oct1_t), target :: grid(noct)
  real(c_double), target :: uold(ttd, nvar, noct), unew(ttd, nvar, noct)
  integer(c_int), target :: ckey_max(nlevel+1), box_min(3, nlevel+1), box_max(3, nlevel+1)
  integer(c_int64_t), target :: key_off(nlevel+1)
  integer(c_int) :: periodic(3), bound_type(6), bmin(3), bmax(3)
  real(c_double) :: gamma, courant, smallr, smallc, dx, t, tout, dt, dtold, trem, dtest, g(3)
  real(c_double) :: mass, ekin, eint, emag, dtc, x, rho, pr
  integer :: i, l, ind, nstep, nrem, nlanding, u
  character(len=512) :: fname

  call get_command_argument(1, fname)
  gamma = 1.4d0; courant = 0.8d0; smallr = 1d-10; smallc = 1d-10
  tout = 0.2d0; nlanding = 3; g = 0d0
  dx = 1d0 / 2**nlevel

  ! level parameters (key_off recursion of metal_runner.f90:26-30)
  do l = 1, nlevel + 1
     ckey_max(l) = 2**(l-1)
     box_min(:, l) = 0
     box_max(:, l) = 1
     box_max(1, l) = ckey_max(l)
  end do
  key_off(1) = 1
  do l = 2, nlevel + 1
     key_off(l) = key_off(l-1) + ckey_max(l-1)
  end do
  periodic = (/ 0, 1, 1 /)
  bound_type = 2

  ! mesh and initial condition (namelist/tube1d.nml:31-40 regions)
  do i = 1, noct
     grid(i)%hkey = 0
     grid(i)%ckey(1) = i - 1
     grid(i)%refined = 0
     grid(i)%lev = nlevel
     do ind = 1, ttd
        x = (dble(2*(i-1) + ind - 1) + 0.5d0) * dx
        if (x < 0.5d0) then
           rho = 1d0; pr = 1d0
        else
           rho = 0.125d0; pr = 0.1d0
        end if
        uold(ind, 1, i) = rho
        uold(ind, 2:4, i) = 0d0
        uold(ind, 5, i) = pr / (gamma - 1d0)
     end do
  end do
  unew = uold

  call rh_check(rh_init(), 'rh_init')
  call rh_check(rh_alloc_amr(noct, 0, nvar, ttd, 2*noct), 'rh_alloc_amr')
  call rh_check(rh_upload_level_params(c_loc(ckey_max), c_loc(key_off), c_loc(box_min), c_loc(box_max), &
       periodic, nlevel), 'rh_upload_level_params')
  call rh_check(rh_set_boundaries(bound_type), 'rh_set_boundaries')
  call rh_check(rh_set_grid_device(c_loc(uold), c_loc(unew), c_null_ptr, c_loc(grid), noct, nvar, ttd), &
       'rh_set_grid_device')
  call rh_check(rh_insert_hash_all(1, noct, 2*noct), 'rh_insert_hash_all')
  bmin = box_min(:, nlevel); bmax = box_max(:, nlevel)
  call rh_check(rh_build_nbor(1, noct, 2*noct, ckey_max(nlevel), key_off(nlevel), bmin, bmax, periodic), &
       'rh_build_nbor')

  t = 0d0; dtold = 0d0; nstep = 0
  do while (t < tout .and. nstep < 5000)
     ! newdt_fine: dx/smallc, Courant, output landing
     call rh_check(rh_cmpdt(1, noct, dx, gamma, smallr, smallc**2, courant, g, mass, ekin, eint, emag, dtc), &
          'rh_cmpdt')
     dt = min(dx / smallc, dtc)
     trem = tout - t
     if (trem > 0d0) then
        dtest = dt
        if (dtold > 0d0) dtest = min(dtest, dtold)
        if (trem < nlanding * dtest) then
           nrem = max(1, ceiling(trem / dtest - 1d-6))
           dt = trem / dble(nrem)
           if (nrem == 1) dt = trem + 1d-10 * max(trem, abs(tout))
        end if
     end if
     call rh_check(rh_set_unew(1, noct), 'rh_set_unew')
     call rh_check(rh_godunov(1, noct, noct, nlevel, nlevel, nlevel, gamma, smallr, smallc**2, dt, dx, &
          courant * dx / dt, 1, 3, g), 'rh_godunov')
     call rh_check(rh_set_uold(1, noct), 'rh_set_uold')
     t = t + dt
     dtold = dt
     nstep = nstep + 1
  end do
  call rh_check(rh_transfer_grid_host(c_loc(uold), c_null_ptr, noct, nvar, ttd), 'rh_transfer_grid_host')

  open(newunit=u, file=trim(fname), access='stream', form='unformatted', status='replace')
  write(u) nstep, t, uold
  close(u)
  write(*,'(A,I5,A,ES23.15)') 'sod_driver: nstep=', nstep, ' t=', t
end program sod_driver
