! Fortran binding of the gfx950 hydro library (include/ramses2_hip.h).
!
! The counterpart of metal/metal_interface.f90 for the HIP backend: the same
! entry points with real(c_double) in place of the Metal backend's
! real(c_float), each returning an integer(c_int) status (0 = success; the
! message of a failure is rh_last_error()).  A RAMSES build adds this module
! and calls it from an "#elif defined(_HIP)" branch at the call sites listed
! in include/ramses2_hip.h.
module hip_interface
  use iso_c_binding
  implicit none

  interface

    integer(c_int) function rh_init() bind(C, name="rh_init")
      import c_int
    end function rh_init

    integer(c_int) function rh_alloc_amr(ngridmax, ncachemax, nvar, twotondim, hash_size) &
         bind(C, name="rh_alloc_amr")
      import c_int
      integer(c_int), value :: ngridmax, ncachemax, nvar, twotondim, hash_size
    end function rh_alloc_amr

    integer(c_int) function rh_upload_level_params(ckey_max, key_off, box_ckey_min, box_ckey_max, &
         periodic, nlevelmax) bind(C, name="rh_upload_level_params")
      import c_int, c_ptr
      type(c_ptr), value :: ckey_max, key_off, box_ckey_min, box_ckey_max
      integer(c_int), intent(in) :: periodic(3)
      integer(c_int), value :: nlevelmax
    end function rh_upload_level_params

    integer(c_int) function rh_set_boundaries(bound_type) bind(C, name="rh_set_boundaries")
      import c_int
      integer(c_int), intent(in) :: bound_type(6)
    end function rh_set_boundaries

    integer(c_int) function rh_set_grid_device(uold, unew, bold, grid, ngridmax, nvar, twotondim) &
         bind(C, name="rh_set_grid_device")
      import c_int, c_ptr
      type(c_ptr), value :: uold, unew, bold, grid
      integer(c_int), value :: ngridmax, nvar, twotondim
    end function rh_set_grid_device

    integer(c_int) function rh_transfer_grid_host(uold, bold, ngridmax, nvar, twotondim) &
         bind(C, name="rh_transfer_grid_host")
      import c_int, c_ptr
      type(c_ptr), value :: uold, bold
      integer(c_int), value :: ngridmax, nvar, twotondim
    end function rh_transfer_grid_host

    integer(c_int) function rh_insert_hash_all(head_idx, num_octs, hash_size) bind(C, name="rh_insert_hash_all")
      import c_int
      integer(c_int), value :: head_idx, num_octs, hash_size
    end function rh_insert_hash_all

    integer(c_int) function rh_build_nbor(head_idx, num_octs, hash_size, ckey_max_l, key_off_l, &
         box_min, box_max, periodic) bind(C, name="rh_build_nbor")
      import c_int, c_long
      integer(c_int), value :: head_idx, num_octs, hash_size, ckey_max_l
      integer(c_long), value :: key_off_l
      integer(c_int), intent(in) :: box_min(3), box_max(3), periodic(3)
    end function rh_build_nbor

    integer(c_int) function rh_set_unew(head_idx, num_octs) bind(C, name="rh_set_unew")
      import c_int
      integer(c_int), value :: head_idx, num_octs
    end function rh_set_unew

    integer(c_int) function rh_set_uold(head_idx, num_octs) bind(C, name="rh_set_uold")
      import c_int
      integer(c_int), value :: head_idx, num_octs
    end function rh_set_uold

    integer(c_int) function rh_cmpdt(head_idx, num_octs, dx, gamma, smallr, smallc2, courant_factor, &
         constant_gravity, mass, ekin, eint, emag, dt) bind(C, name="rh_cmpdt")
      import c_int, c_double
      integer(c_int), value :: head_idx, num_octs
      real(c_double), value :: dx, gamma, smallr, smallc2, courant_factor
      real(c_double), intent(in) :: constant_gravity(3)
      real(c_double), intent(out) :: mass, ekin, eint, emag, dt
    end function rh_cmpdt

    integer(c_int) function rh_godunov(head_idx, num_subgrids, ngridmax, ilevel, levelmin, levelmax, &
         gamma, smallr, smallc2, dt, dx, smax_cap, slope, riemann, constant_gravity) bind(C, name="rh_godunov")
      import c_int, c_double
      integer(c_int), value :: head_idx, num_subgrids, ngridmax, ilevel, levelmin, levelmax
      real(c_double), value :: gamma, smallr, smallc2, dt, dx, smax_cap
      integer(c_int), value :: slope, riemann
      real(c_double), intent(in) :: constant_gravity(3)
    end function rh_godunov

    ! refinement flags (mtl_upload_flag1 / mtl_reset_flag1 / mtl_count_flag1 /
    ! mtl_hydro_flag, metal/metal_interface.f90:59-63,200-241)
    integer(c_int) function rh_upload_flag1(flag1, ngridmax) bind(C, name="rh_upload_flag1")
      import c_int
      integer(c_int), value :: ngridmax
      integer(c_int), intent(in) :: flag1(*)
    end function rh_upload_flag1

    integer(c_int) function rh_transfer_flag1_host(flag1, ngridmax) bind(C, name="rh_transfer_flag1_host")
      import c_int
      integer(c_int), value :: ngridmax
      integer(c_int), intent(out) :: flag1(*)
    end function rh_transfer_flag1_host

    integer(c_int) function rh_reset_flag1(head_idx, num_octs) bind(C, name="rh_reset_flag1")
      import c_int
      integer(c_int), value :: head_idx, num_octs
    end function rh_reset_flag1

    integer(c_int) function rh_count_flag1(head_idx, num_octs, count) bind(C, name="rh_count_flag1")
      import c_int
      integer(c_int), value :: head_idx, num_octs
      integer(c_int), intent(out) :: count
    end function rh_count_flag1

    integer(c_int) function rh_hydro_flag(head_idx, num_octs, gamma, smallr, smallc2, err_grad_d, err_grad_p, &
         floor_d, floor_p) bind(C, name="rh_hydro_flag")
      import c_int, c_double
      integer(c_int), value :: head_idx, num_octs
      real(c_double), value :: gamma, smallr, smallc2, err_grad_d, err_grad_p, floor_d, floor_p
    end function rh_hydro_flag

    integer(c_int) function rh_device_sync() bind(C, name="rh_device_sync")
      import c_int
    end function rh_device_sync

    type(c_ptr) function rh_last_error() bind(C, name="rh_last_error")
      import c_ptr
    end function rh_last_error

    integer(c_int) function rh_upload(head_idx, num_octs, internal_energy, gamma, smallr, smallc2) &
         bind(C, name="rh_upload")
      import c_int, c_double
      integer(c_int), value :: head_idx, num_octs, internal_energy
      real(c_double), value :: gamma, smallr, smallc2
    end function rh_upload

    integer(c_int) function rh_set_interpolation(interpol_var, interpol_type) bind(C, name="rh_set_interpolation")
      import c_int
      integer(c_int), value :: interpol_var, interpol_type
    end function rh_set_interpolation

    ! ---- extensions (include/ramses2_hip_ext.h) ----
    integer(c_int) function rh_set_precision(precision) bind(C, name="rh_set_precision")
      import c_int
      integer(c_int), value :: precision
    end function rh_set_precision

    integer(c_int) function rh_amr_prepare(levelmin, levelmax, head, noct, boxlen, gamma, smallr, smallc2, &
         courant_factor, constant_gravity, slope, riemann) bind(C, name="rh_amr_prepare")
      import c_int, c_double, c_ptr
      integer(c_int), value :: levelmin, levelmax, slope, riemann
      type(c_ptr), value :: head, noct
      real(c_double), value :: boxlen, gamma, smallr, smallc2, courant_factor
      real(c_double), intent(in) :: constant_gravity(3)
    end function rh_amr_prepare

    integer(c_int) function rh_amr_run(nsteps, precision, use_graph, dt_last, t_total, sums, ms_per_step) &
         bind(C, name="rh_amr_run")
      import c_int, c_double, c_float
      integer(c_int), value :: nsteps, precision, use_graph
      real(c_double), intent(out) :: dt_last, t_total, sums(4)
      real(c_float), intent(out) :: ms_per_step
    end function rh_amr_run

    integer(c_int) function rh_set_brick_shape(bx, by, bz) bind(C, name="rh_set_brick_shape")
      import c_int
      integer(c_int), value :: bx, by, bz
    end function rh_set_brick_shape

  end interface

contains

  ! Message of the last failed call, as a Fortran string (the runner prints it
  ! before error stop, as metal_runner.f90 does for Metal failures).
  function rh_error_message() result(msg)
    character(len=256) :: msg
    character(kind=c_char), pointer :: cmsg(:)
    integer :: i
    msg = ' '
    call c_f_pointer(rh_last_error(), cmsg, [256])
    do i = 1, 256
       if (cmsg(i) == c_null_char) exit
       msg(i:i) = cmsg(i)
    end do
  end function rh_error_message

  ! error stop on a nonzero status, naming the entry point
  subroutine rh_check(status, where)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: where
    if (status /= 0) then
       write(*,'(A,A,I3,A,A)') where, ' failed, status ', status, ': ', trim(rh_error_message())
       error stop 'HIP hydro backend failure'
    end if
  end subroutine rh_check

end module hip_interface
