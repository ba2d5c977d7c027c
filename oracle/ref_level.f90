! Test harness around the reference's own level routines.
!
! TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile (target reflev)
! together with the reference sources hydro/godunov_fine.f90,
! hydro/courant_fine.f90, hydro/upload.f90, hydro/umuscl.f90,
! hydro/godunov_utils.f90, hydro/interpol_hydro.f90 and the modules they use
! (amr/cache.f90, amr/nbors_utils.f90, amr/boundaries.f90, amr/hash.f90,
! mdl1/mdl.f90, ...), taken unchanged from the reference tree with
! -DWITHOUTMPI, into oracle/_ref/libramses_reflev{1,3}d.so.
!
! The routines below build one mesh in the reference's own types
! (ramses_t/mesh_t, the grid_dict hash filled with hash_setp, the hydro
! kernel workspace of init_hydro_kernel) from plain C arrays, following the
! allocations of amr/init_amr.f90:95-457, and then call the reference's
! godunov_fine, courant_fine, upload_fine, set_unew and set_uold unchanged.
! Tests compare the C restatement in oracle/ramses_oracle.c with them bit
! for bit.  Nothing here is part of the product.
module ref_level
  use iso_c_binding
  use amr_parameters, only: ndim, twotondim, twondim, nhilbert, nsubgrid
  use hydro_parameters, only: nvar
  use amr_commons, only: run_t, global_t, mesh_t, set_hydro_parameters
  use ramses_commons, only: ramses_t
  use mdl_module, only: mdl_t
  use hash, only: init_empty_hash, hash_setp
  use godunov_fine_module, only: godunov_fine, set_unew, set_uold
  use courant_fine_module, only: courant_fine
  use upload_module, only: upload_fine
  use hydro_flag_module, only: hydro_flag
  use cache, only: open_cache
  use nbors_utils, only: get_grid
  use boundaries, only: init_bound_refine
  implicit none

  type(ramses_t), target, save :: s
  type(mesh_t), target, save :: mesh
  type(mdl_t), target, save :: mdl
  logical, save :: ready = .false.

contains

  function reflev_ndim() bind(C, name="reflev_ndim") result(n)
    integer(c_int) :: n
    n = ndim
  end function reflev_ndim

  ! Build the mesh.  ckey(3,ngrid), lev(ngrid), refined(8,ngrid) (0/1),
  ! uold/unew(twotondim,nvar,ngrid); head/noct per level 1..nlevelmax
  ! (1-based oct indices).  The computational box is, at level l >= box_lev,
  ! [box_min*2^(l-box_lev), box_max*2^(l-box_lev)) per dimension (box_lev = 0:
  ! the whole level).  Non-periodic dimensions get one boundary region per
  ! side: the slab outside the box, full width in the other dimensions
  ! (only one dimension may be non-periodic in 3D: at edges and corners the
  ! reference's ghost depends on which ghosts its cache already holds),
  ! bound_type(side,dim) (1 reflexive, 2 zero gradient), its reference oct
  ! one step towards the box (bound_dir = dim, bound_shift = +1 low, -1 high).
  subroutine reflev_setup(ngrid, ckey, lev, refined, uold, unew, levelmin, nlevelmax, head, noct, &
       & boxlen, periodic, bound_type, box_lev, box_min, box_max, interpol_var, interpol_type, grav, &
       & gamma, smallr, smallc, courant_factor, slope_type, riemann, difmag) bind(C, name="reflev_setup")
    integer(c_int), value :: ngrid, levelmin, nlevelmax, box_lev, interpol_var, interpol_type
    integer(c_int), value :: slope_type, riemann
    integer(c_int32_t), dimension(3, ngrid) :: ckey
    integer(c_int32_t), dimension(ngrid) :: lev
    integer(c_int8_t), dimension(8, ngrid) :: refined
    real(c_double), dimension(twotondim, nvar, ngrid) :: uold, unew
    integer(c_int32_t), dimension(nlevelmax) :: head, noct
    real(c_double), value :: boxlen, gamma, smallr, smallc, courant_factor, difmag
    integer(c_int), dimension(3) :: periodic, box_min, box_max
    integer(c_int), dimension(2, 3) :: bound_type
    real(c_double), dimension(3) :: grav
    integer :: i, l, d, ib, sd, blev, bmin(3), bmax(3)
    integer(kind=8), dimension(0:ndim) :: key

    if (ready) call reflev_free()
    s%r%verbose = .false.
    s%r%levelmin = levelmin
    s%r%nlevelmax = nlevelmax
    s%r%ngridmax = ngrid
    s%r%ncachemax = 10000
    s%r%boxlen = boxlen
    s%r%gamma = gamma
    s%r%smallr = smallr
    s%r%smallc = smallc
    s%r%courant_factor = courant_factor
    s%r%slope_type = slope_type
    s%r%riemann = riemann
    s%r%difmag = difmag
    s%r%interpol_var = interpol_var
    s%r%interpol_type = interpol_type
    s%r%induction = .false.
    s%r%constant_gravity = grav
    s%g%myid = 1
    s%g%ncpu = 1
    call set_hydro_parameters(s%r, s%h_params)

    ! box at level blev
    if (box_lev <= 0) then
       blev = 1
       bmin = 0
       bmax = 1
    else
       blev = box_lev
       bmin = box_min
       bmax = box_max
    end if
    s%r%bound_levelmin = blev
    s%r%nbound = 0
    do d = 1, ndim
       s%r%periodic(d) = periodic(d) /= 0
       if (periodic(d) == 0) then
          do sd = 1, 2
             s%r%nbound = s%r%nbound + 1
             ib = s%r%nbound
             s%r%bound_type(ib) = bound_type(sd, d)
             s%r%bound_dir(ib) = d
             s%r%bound_shift(ib) = merge(1, -1, sd == 1)
          end do
       end if
    end do

    ! mesh arrays as amr/init_amr.f90:126-226
    mesh%ngridmax = ngrid
    mesh%ncachemax = max(s%r%ncachemax, 10000)
    allocate(mesh%grid(1:mesh%ngridmax + mesh%ncachemax))
    do i = 1, mesh%ngridmax + mesh%ncachemax
       mesh%grid(i)%lev = 0
       mesh%grid(i)%hkey = 0
       mesh%grid(i)%ckey = 0
       mesh%grid(i)%refined = .false.
    end do
    allocate(mesh%flag1(1:twotondim, 1:mesh%ngridmax + mesh%ncachemax))
    allocate(mesh%flag2(1:twotondim, 1:mesh%ngridmax + mesh%ncachemax))
    allocate(mesh%uold(1:twotondim, 1:nvar, 1:mesh%ngridmax + mesh%ncachemax))
    allocate(mesh%unew(1:twotondim, 1:nvar, 1:mesh%ngridmax + mesh%ncachemax))
    mesh%uold = 0.0d0
    mesh%unew = 0.0d0
    allocate(mesh%dirty(1:mesh%ncachemax), mesh%locked(1:mesh%ncachemax), mesh%occupied(1:mesh%ncachemax))
    allocate(mesh%parent_cpu(1:mesh%ncachemax), mesh%ghost_parent_grid(1:mesh%ncachemax))
    allocate(mesh%ghost_parent_cell(1:mesh%ncachemax))
    mesh%dirty = .false.
    mesh%locked = .false.
    mesh%occupied = .false.
    mesh%ghost_parent_grid = 0
    mesh%ghost_parent_cell = 0
    mesh%free_cache = 1; mesh%ncache = 0; mesh%nlocked = 0; mesh%nlocked_max = 0
    allocate(mesh%lev_null(1:mesh%ncachemax), mesh%ckey_null(1:ndim, 1:mesh%ncachemax))
    allocate(mesh%occupied_null(1:mesh%ncachemax))
    mesh%occupied_null = .false.
    mesh%free_null = 1; mesh%nnull = 0
    call init_empty_hash(mesh%grid_dict, 2*(mesh%ngridmax + mesh%ncachemax), 'simple')
    allocate(mesh%ckey_max(1:nlevelmax + 1))
    allocate(mesh%box_ckey_min(1:3, 1:nlevelmax + 1), mesh%box_ckey_max(1:3, 1:nlevelmax + 1))
    mesh%box_ckey_min = 0
    mesh%box_ckey_max = 1
    if (s%r%nbound > 0) then
       allocate(mesh%bound_ckey_min(1:3, 1:s%r%nbound, 1:nlevelmax + 1))
       allocate(mesh%bound_ckey_max(1:3, 1:s%r%nbound, 1:nlevelmax + 1))
    end if
    do l = 1, nlevelmax + 1
       mesh%ckey_max(l) = 2**(l - 1)
    end do
    do l = blev, nlevelmax + 1
       do d = 1, 3
          mesh%box_ckey_min(d, l) = bmin(d)*2**(l - blev)
          mesh%box_ckey_max(d, l) = bmax(d)*2**(l - blev)
       end do
       do ib = 1, s%r%nbound
          ! slab outside the box on one side of dimension bound_dir, wide
          ! enough to hold the stencil keys that leave the level (-1, 2^(l-1))
          mesh%bound_ckey_min(1:3, ib, l) = -4
          mesh%bound_ckey_max(1:3, ib, l) = mesh%ckey_max(l) + 4
          d = s%r%bound_dir(ib)
          if (s%r%bound_shift(ib) == 1) then
             mesh%bound_ckey_max(d, ib, l) = mesh%box_ckey_min(d, l)
          else
             mesh%bound_ckey_min(d, ib, l) = mesh%box_ckey_max(d, l)
          end if
       end do
    end do
    allocate(mesh%head(1:nlevelmax), mesh%tail(1:nlevelmax), mesh%noct(1:nlevelmax))
    allocate(mesh%noct_min(1:nlevelmax), mesh%noct_max(1:nlevelmax), mesh%noct_tot(1:nlevelmax))
    do l = 1, nlevelmax
       mesh%head(l) = head(l)
       mesh%noct(l) = noct(l)
       mesh%tail(l) = head(l) + noct(l) - 1
       mesh%noct_tot(l) = noct(l)
       mesh%noct_min(l) = noct(l)
       mesh%noct_max(l) = noct(l)
    end do

    ! octs, state and hash
    do i = 1, ngrid
       mesh%grid(i)%lev = lev(i)
       mesh%grid(i)%ckey(1:ndim) = ckey(1:ndim, i)
       do d = 1, twotondim
          mesh%grid(i)%refined(d) = refined(d, i) /= 0
       end do
       mesh%uold(:, :, i) = uold(:, :, i)
       mesh%unew(:, :, i) = unew(:, :, i)
       key(0) = lev(i)
       key(1:ndim) = ckey(1:ndim, i)
       call hash_setp(mesh%grid_dict, key, i)
    end do
    call mesh%hydro_w%kernel%init(2*nsubgrid)

    s%m => mesh
    s%mdl => mdl
    mdl%m => mesh
    ready = .true.
  end subroutine reflev_setup

  ! godunov_fine (hydro/godunov_fine.f90:43-79) with g%dtnew(ilevel) = dt
  subroutine reflev_godunov(ilevel, dt) bind(C, name="reflev_godunov")
    integer(c_int), value :: ilevel
    real(c_double), value :: dt
    s%g%dtnew(ilevel) = dt
    call godunov_fine(s, ilevel)
  end subroutine reflev_godunov

  ! courant_fine (hydro/courant_fine.f90:57-168); out = mass, ekin, eint, emag, dt
  subroutine reflev_courant(ilevel, out) bind(C, name="reflev_courant")
    integer(c_int), value :: ilevel
    real(c_double), dimension(5) :: out
    real(kind=8) :: mass, ekin, eint, emag, dt
    call courant_fine(s%r, s%g, mesh, ilevel, mass, ekin, eint, emag, dt)
    out = (/mass, ekin, eint, emag, dt/)
  end subroutine reflev_courant

  ! upload_fine (hydro/upload.f90:74-283)
  subroutine reflev_upload(ilevel) bind(C, name="reflev_upload")
    integer(c_int), value :: ilevel
    call upload_fine(s, ilevel)
  end subroutine reflev_upload

  subroutine reflev_set_unew(ilevel) bind(C, name="reflev_set_unew")
    integer(c_int), value :: ilevel
    call set_unew(s%r, s%g, mesh, ilevel)
  end subroutine reflev_set_unew

  subroutine reflev_set_uold(ilevel) bind(C, name="reflev_set_uold")
    integer(c_int), value :: ilevel
    call set_uold(s%r, s%g, mesh, ilevel)
  end subroutine reflev_set_uold

  ! get_grid (amr/nbors_utils.f90:298-667) of the key (lev, key(1:ndim)) with
  ! a read-only cache whose boundary initialiser is init_bound_refine
  ! (amr/boundaries.f90:39-523), as godunov_fine opens it.  Returns the oct
  ! index (a real oct, or a boundary ghost built in the cache region,
  ! > ngridmax) or 0, and that oct's state and refined map.  The cache is
  ! left open (closing it would destage the ghosts, which needs
  ! amr/task_manager.f90): call reflev_setup again before the next godunov.
  function reflev_get_grid(lev, key, u, refined) bind(C, name="reflev_get_grid") result(child)
    integer(c_int), value :: lev
    integer(c_int), dimension(3) :: key
    real(c_double), dimension(twotondim, nvar) :: u
    integer(c_int8_t), dimension(twotondim) :: refined
    integer(c_int) :: child
    integer(kind=8), dimension(0:ndim) :: hk
    integer :: ic, d
    if (.not. mdl%cache_opened) call open_cache(mdl, mesh, pack_size=1, bound=init_bound_refine)
    hk(0) = lev
    hk(1:ndim) = key(1:ndim)
    call get_grid(s, hk, ic, flush_cache=.false., fetch_cache=.true.)
    child = ic
    u = 0.0d0
    refined = 0
    if (ic > 0) then
       u = mesh%uold(:, :, ic)
       do d = 1, twotondim
          if (mesh%grid(ic)%refined(d)) refined(d) = 1
       end do
    end if
  end function reflev_get_grid

  ! hydro_flag (hydro/hydro_flag.f90:7-159, the gradient criteria of
  ! hydro_refine, hydro/godunov_utils.f90:6-227) on the level's cells, flag1
  ! cleared first; flag1 of the first ngrid octs is copied out.  Periodic
  ! meshes only (the routine closes its cache, see reflev_godunov).
  subroutine reflev_hydro_flag(ilevel, err_grad_d, err_grad_p, floor_d, floor_p, ngrid, flag1) &
       & bind(C, name="reflev_hydro_flag")
    integer(c_int), value :: ilevel, ngrid
    real(c_double), value :: err_grad_d, err_grad_p, floor_d, floor_p
    integer(c_int32_t), dimension(twotondim, ngrid) :: flag1
    s%r%err_grad_d = err_grad_d
    s%r%err_grad_p = err_grad_p
    s%r%floor_d = floor_d
    s%r%floor_p = floor_p
    s%r%m_refine = -1.0
    s%r%poisson = .false.
    mesh%flag1 = 0
    s%g%nflag = 0
    call hydro_flag(s, ilevel)
    flag1 = int(mesh%flag1(:, 1:ngrid), c_int32_t)
  end subroutine reflev_hydro_flag

  ! Copy the first ngrid octs of uold and unew out.
  subroutine reflev_get(ngrid, uold, unew) bind(C, name="reflev_get")
    integer(c_int), value :: ngrid
    real(c_double), dimension(twotondim, nvar, ngrid) :: uold, unew
    uold = mesh%uold(:, :, 1:ngrid)
    unew = mesh%unew(:, :, 1:ngrid)
  end subroutine reflev_get

  subroutine reflev_free() bind(C, name="reflev_free")
    if (.not. ready) return
    mdl%cache_opened = .false.
    deallocate(mesh%grid, mesh%flag1, mesh%flag2, mesh%uold, mesh%unew)
    deallocate(mesh%dirty, mesh%locked, mesh%occupied, mesh%parent_cpu)
    deallocate(mesh%ghost_parent_grid, mesh%ghost_parent_cell)
    deallocate(mesh%lev_null, mesh%ckey_null, mesh%occupied_null)
    deallocate(mesh%ckey_max, mesh%box_ckey_min, mesh%box_ckey_max)
    if (allocated(mesh%bound_ckey_min)) deallocate(mesh%bound_ckey_min, mesh%bound_ckey_max)
    deallocate(mesh%head, mesh%tail, mesh%noct, mesh%noct_min, mesh%noct_max, mesh%noct_tot)
    deallocate(mesh%hydro_w%kernel%uloc, mesh%hydro_w%kernel%gloc, mesh%hydro_w%kernel%qloc)
    deallocate(mesh%hydro_w%kernel%cloc, mesh%hydro_w%kernel%okloc, mesh%hydro_w%kernel%dq)
    deallocate(mesh%hydro_w%kernel%qm, mesh%hydro_w%kernel%qp, mesh%hydro_w%kernel%fx)
    deallocate(mesh%hydro_w%kernel%tx, mesh%hydro_w%kernel%flux, mesh%hydro_w%kernel%tmp)
    deallocate(mesh%hydro_w%kernel%divu, mesh%hydro_w%kernel%childloc, mesh%hydro_w%kernel%gridloc)
    deallocate(mesh%hydro_w%kernel%cellloc, mesh%hydro_w%kernel%nborloc)
    call reset_hash()
    ready = .false.
  end subroutine reflev_free

  subroutine reset_hash()
    if (allocated(mesh%grid_dict%data)) deallocate(mesh%grid_dict%data)
    if (allocated(mesh%grid_dict%next_free)) deallocate(mesh%grid_dict%next_free)
  end subroutine reset_hash

end module ref_level
