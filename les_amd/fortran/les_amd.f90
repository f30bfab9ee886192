! les_amd.f90 -- Fortran (iso_c_binding) side of the C ABI in include/les_amd.h.
!
! module les_amd      : bind(C) interfaces, one per C entry point (device
!                        pointers as type(c_ptr)), plus les_check.
! module les_amd_host : host-array wrappers with the reference routines'
!                        argument lists (filt_da.f90, ddx.f90, convec.f90 ...).
!                        Each keeps a device mirror per host array (les_mirror),
!                        pushes the inputs, runs the HIP routine and pulls the
!                        outputs, so a reference routine body can be replaced by
!                        one call.  The whole-step path (les_step) keeps the
!                        state on the GPU instead and is the fast path.
module les_amd
  use iso_c_binding
  implicit none

  type, bind(C) :: les_grid
    integer(c_int) :: nx, ny, nz
    real(c_double) :: Lx, Ly, Lz, dt, z0, force
    integer(c_int) :: inilag, initu, cs_count
  end type les_grid

  ! les_scalar_params: param.f90:53,72-90,139-146,174 and bottombc.f90:26-63
  type, bind(C) :: les_scalar_params
    integer(c_int) :: enable, model, initsc, passive
    real(c_double) :: Pr, z_i, u_star, T_scale, q_scale, g, dTdz_top, dqdz_top
  end type les_scalar_params

  interface
    integer(c_int) function les_create(g, ctx) bind(C, name='les_create')
      import :: c_int, c_ptr, les_grid
      type(les_grid), intent(in) :: g
      type(c_ptr), intent(out) :: ctx
    end function
    subroutine les_destroy(ctx) bind(C, name='les_destroy')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine
    type(c_ptr) function les_last_error(ctx) bind(C, name='les_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function les_get_grid(ctx, g) bind(C, name='les_get_grid')
      import :: c_int, c_ptr, les_grid
      type(c_ptr), value :: ctx
      type(les_grid), intent(out) :: g
    end function
    integer(c_int) function les_sync(ctx) bind(C, name='les_sync')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function les_set_time(ctx, jt, jt_total) bind(C, name='les_set_time')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: jt, jt_total
    end function
    ! --- routine-level (device pointers) ---
    integer(c_int) function les_filt_da(ctx, f, dfdx, dfdy, nlev) bind(C, name='les_filt_da')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, f, dfdx, dfdy
      integer(c_int), value :: nlev
    end function
    integer(c_int) function les_ddx(ctx, dfdx, f, nlev) bind(C, name='les_ddx')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, dfdx, f
      integer(c_int), value :: nlev
    end function
    integer(c_int) function les_ddy(ctx, dfdy, f, nlev) bind(C, name='les_ddy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, dfdy, f
      integer(c_int), value :: nlev
    end function
    integer(c_int) function les_ddxy(ctx, dfdx, dfdy, f, nlev) bind(C, name='les_ddxy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, dfdx, dfdy, f
      integer(c_int), value :: nlev
    end function
    integer(c_int) function les_test_filter(ctx, plane, g) bind(C, name='les_test_filter')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, plane, g
    end function
    integer(c_int) function les_dealias1(ctx, u, u_big, nlev) bind(C, name='les_dealias1')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, u, u_big
      integer(c_int), value :: nlev
    end function
    integer(c_int) function les_dealias2(ctx, u, u_big, nlev) bind(C, name='les_dealias2')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, u, u_big
      integer(c_int), value :: nlev
    end function
    integer(c_int) function les_ddz_uv(ctx, dfdz, f) bind(C, name='les_ddz_uv')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, dfdz, f
    end function
    integer(c_int) function les_ddz_w(ctx, dfdz, f) bind(C, name='les_ddz_w')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, dfdz, f
    end function
    integer(c_int) function les_convec(ctx, u, v, w, dudy, dudz, dvdx, dvdz, dwdx, dwdy, cx, cy, cz) &
        bind(C, name='les_convec')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, u, v, w, dudy, dudz, dvdx, dvdz, dwdx, dwdy, cx, cy, cz
    end function
    integer(c_int) function les_divstress_uv(ctx, divt, tx, ty, tz) bind(C, name='les_divstress_uv')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, divt, tx, ty, tz
    end function
    integer(c_int) function les_divstress_w(ctx, divt, tx, ty, tz) bind(C, name='les_divstress_w')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, divt, tx, ty, tz
    end function
    integer(c_int) function les_press_stag_array(ctx, u, v, w, divtz, p, dpdx, dpdy) &
        bind(C, name='les_press_stag_array')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, u, v, w, divtz, p, dpdx, dpdy
    end function
    integer(c_int) function les_project(ctx, u, v, w, dpdx, dpdy, dpdz) bind(C, name='les_project')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, u, v, w, dpdx, dpdy, dpdz
    end function
    integer(c_int) function les_rmsdiv(ctx, dudx, dvdy, dwdz, rms) bind(C, name='les_rmsdiv')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, dudx, dvdy, dwdz
      real(c_double), intent(out) :: rms
    end function
    ! --- mirrors ---
    type(c_ptr) function les_mirror(ctx, host, bytes) bind(C, name='les_mirror')
      import :: c_ptr, c_int64_t
      type(c_ptr), value :: ctx, host
      integer(c_int64_t), value :: bytes
    end function
    integer(c_int) function les_mirror_push(ctx, host, bytes) bind(C, name='les_mirror_push')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, host
      integer(c_int64_t), value :: bytes
    end function
    integer(c_int) function les_mirror_pull(ctx, host, bytes) bind(C, name='les_mirror_pull')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, host
      integer(c_int64_t), value :: bytes
    end function
    integer(c_int) function les_memcpy_d2h(ctx, dst, src, bytes) bind(C, name='les_memcpy_d2h')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, dst, src
      integer(c_int64_t), value :: bytes
    end function
    ! --- whole-step state ---
    integer(c_int) function les_state_alloc(ctx, precision) bind(C, name='les_state_alloc')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: precision
    end function
    type(c_ptr) function les_state_field(ctx, name, nlev, elsize) bind(C, name='les_state_field')
      import :: c_ptr, c_char, c_int
      type(c_ptr), value :: ctx
      character(kind=c_char), dimension(*), intent(in) :: name
      integer(c_int), intent(out) :: nlev, elsize
    end function
    integer(c_int) function les_state_init_ic(ctx, seed, amp, ustar, kappa_ic) bind(C, name='les_state_init_ic')
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: seed
      real(c_double), value :: amp, ustar, kappa_ic
    end function
    integer(c_int) function les_step(ctx, nsteps) bind(C, name='les_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
    end function
    integer(c_int) function les_set_graph(ctx, enable) bind(C, name='les_set_graph')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable
    end function
    ! --- SGS model selection (param.f90 'model': 5 or 3) and scaledep_dynamic.f90 ---
    integer(c_int) function les_set_sgs_model(ctx, model) bind(C, name='les_set_sgs_model')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: model
    end function
    integer(c_int) function les_scaledep_dynamic(ctx, S, u, v, w, Cs_opt2, beta) &
         bind(C, name='les_scaledep_dynamic')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), intent(in) :: S(6)
      type(c_ptr), value :: u, v, w, Cs_opt2, beta
    end function
    ! --- scalars: theta and q (S_FLAG), convective surface fluxes ---
    integer(c_int) function les_set_scalars(ctx, p, wt_s, wq_s) bind(C, name='les_set_scalars')
      import :: c_int, c_ptr, les_scalar_params
      type(c_ptr), value :: ctx
      type(les_scalar_params), intent(in) :: p
      type(c_ptr), value :: wt_s, wq_s
    end function
    integer(c_int) function les_state_init_scalar_ic(ctx, seed, T_init, z_inv, inv_strength, amp_K, q_init) &
         bind(C, name='les_state_init_scalar_ic')
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: seed
      real(c_double), value :: T_init, z_inv, inv_strength, amp_K, q_init
    end function
    integer(c_int) function les_scalar_ref(ctx, theta_ref, q_ref) bind(C, name='les_scalar_ref')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: theta_ref, q_ref
    end function
  end interface

  ! context used by the host wrappers (set by les_init_global)
  type(c_ptr), save :: les_ctx_global = c_null_ptr

contains

  subroutine les_check(rc, where)
    ! the reference stops on fatal conditions (tridag_array.f90:34-38,
    ! wallstress.f90:62-65); non-zero status from the library does the same
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == 0) return
    call c_f_pointer(les_last_error(les_ctx_global), msg, [512])
    n = 0
    do while (n < 512)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    end do
    write (*, '(a,a,a,512a)') 'les_amd: ', where, ': ', msg(1:n)
    error stop 1
  end subroutine les_check

  subroutine les_init_global(g)
    type(les_grid), intent(in) :: g
    call les_check(les_create(g, les_ctx_global), 'les_create')
  end subroutine les_init_global

end module les_amd


module les_amd_host
  use iso_c_binding
  use les_amd
  implicit none
contains

  integer(c_int64_t) function field_bytes(nlev)
    integer, intent(in) :: nlev
    type(les_grid) :: g
    call les_check(les_get_grid(les_ctx_global, g), 'les_get_grid')
    field_bytes = int(2 * (g%nx / 2 + 1), c_int64_t) * g%ny * nlev * 8_c_int64_t
  end function field_bytes

  type(c_ptr) function pushed(a, nb)
    real(c_double), target, intent(in) :: a(*)
    integer(c_int64_t), intent(in) :: nb
    call les_check(les_mirror_push(les_ctx_global, c_loc(a(1)), nb), 'mirror push')
    pushed = les_mirror(les_ctx_global, c_loc(a(1)), nb)
  end function pushed

  type(c_ptr) function dev(a, nb)
    real(c_double), target, intent(in) :: a(*)
    integer(c_int64_t), intent(in) :: nb
    dev = les_mirror(les_ctx_global, c_loc(a(1)), nb)
  end function dev

  subroutine pull(a, nb)
    real(c_double), target, intent(inout) :: a(*)
    integer(c_int64_t), intent(in) :: nb
    call les_check(les_mirror_pull(les_ctx_global, c_loc(a(1)), nb), 'mirror pull')
  end subroutine pull

  ! filt_da.f90:3-42 : filt_da(f_c, dfdx_c, dfdy_c), levels lbz:nz
  subroutine les_h_filt_da(f, dfdx, dfdy, nlev)
    real(c_double), target :: f(*), dfdx(*), dfdy(*)
    integer, intent(in) :: nlev
    integer(c_int64_t) :: nb
    nb = field_bytes(nlev)
    call les_check(les_filt_da(les_ctx_global, pushed(f, nb), dev(dfdx, nb), dev(dfdy, nb), nlev), 'filt_da')
    call pull(f, nb)
    call pull(dfdx, nb)
    call pull(dfdy, nb)
  end subroutine les_h_filt_da

  ! ddx.f90:1-31 : ddx(dfdx, f)
  subroutine les_h_ddx(dfdx, f, nlev)
    real(c_double), target :: dfdx(*), f(*)
    integer, intent(in) :: nlev
    integer(c_int64_t) :: nb
    nb = field_bytes(nlev)
    call les_check(les_ddx(les_ctx_global, dev(dfdx, nb), pushed(f, nb), nlev), 'ddx')
    call pull(dfdx, nb)
  end subroutine les_h_ddx

  ! ddy.f90:1-31 : ddy(dfdy, f)
  subroutine les_h_ddy(dfdy, f, nlev)
    real(c_double), target :: dfdy(*), f(*)
    integer, intent(in) :: nlev
    integer(c_int64_t) :: nb
    nb = field_bytes(nlev)
    call les_check(les_ddy(les_ctx_global, dev(dfdy, nb), pushed(f, nb), nlev), 'ddy')
    call pull(dfdy, nb)
  end subroutine les_h_ddy

  ! ddz_uv.f90:5-57 / ddz_w.f90:6-51 : (ld,ny,1:nz)
  subroutine les_h_ddz_uv(dfdz, f, nz)
    real(c_double), target :: dfdz(*), f(*)
    integer, intent(in) :: nz
    integer(c_int64_t) :: nb
    nb = field_bytes(nz)
    call les_check(les_ddz_uv(les_ctx_global, pushed(dfdz, nb), pushed(f, nb)), 'ddz_uv')
    call pull(dfdz, nb)
  end subroutine les_h_ddz_uv

  subroutine les_h_ddz_w(dfdz, f, nz)
    real(c_double), target :: dfdz(*), f(*)
    integer, intent(in) :: nz
    integer(c_int64_t) :: nb
    nb = field_bytes(nz)
    call les_check(les_ddz_w(les_ctx_global, pushed(dfdz, nb), pushed(f, nb)), 'ddz_w')
    call pull(dfdz, nb)
  end subroutine les_h_ddz_w

  ! convec.f90:1-248 : convec(cx, cy, cz) with its sim_param inputs explicit
  subroutine les_h_convec(u, v, w, dudy, dudz, dvdx, dvdz, dwdx, dwdy, cx, cy, cz, nz)
    real(c_double), target :: u(*), v(*), w(*), dudy(*), dudz(*), dvdx(*), dvdz(*), dwdx(*), dwdy(*)
    real(c_double), target :: cx(*), cy(*), cz(*)
    integer, intent(in) :: nz
    integer(c_int64_t) :: nb
    nb = field_bytes(nz)
    call les_check(les_convec(les_ctx_global, pushed(u, nb), pushed(v, nb), pushed(w, nb), pushed(dudy, nb), &
         pushed(dudz, nb), pushed(dvdx, nb), pushed(dvdz, nb), pushed(dwdx, nb), pushed(dwdy, nb), &
         dev(cx, nb), dev(cy, nb), dev(cz, nb)), 'convec')
    call pull(cx, nb)
    call pull(cy, nb)
    call pull(cz, nb)
  end subroutine les_h_convec

  ! divstress_uv.f90:1-63 / divstress_w.f90:4-86
  subroutine les_h_divstress_uv(divt, tx, ty, tz, nz)
    real(c_double), target :: divt(*), tx(*), ty(*), tz(*)
    integer, intent(in) :: nz
    integer(c_int64_t) :: nb
    nb = field_bytes(nz)
    call les_check(les_divstress_uv(les_ctx_global, dev(divt, nb), pushed(tx, nb), pushed(ty, nb), &
         pushed(tz, nb)), 'divstress_uv')
    call pull(divt, nb)
  end subroutine les_h_divstress_uv

  subroutine les_h_divstress_w(divt, tx, ty, tz, nz)
    real(c_double), target :: divt(*), tx(*), ty(*), tz(*)
    integer, intent(in) :: nz
    integer(c_int64_t) :: nb
    nb = field_bytes(nz)
    call les_check(les_divstress_w(les_ctx_global, dev(divt, nb), pushed(tx, nb), pushed(ty, nb), &
         pushed(tz, nb)), 'divstress_w')
    call pull(divt, nb)
  end subroutine les_h_divstress_w

  ! press_stag_array.f90:1-294 : press_stag_array(p_hat, dfdx, dfdy) with
  ! sim_param u, v, w, divtz explicit; p is (ld,ny,0:nz)
  subroutine les_h_press_stag_array(u, v, w, divtz, p, dpdx, dpdy, nz)
    real(c_double), target :: u(*), v(*), w(*), divtz(*), p(*), dpdx(*), dpdy(*)
    integer, intent(in) :: nz
    integer(c_int64_t) :: nb, nbp
    nb = field_bytes(nz)
    nbp = field_bytes(nz + 1)
    call les_check(les_press_stag_array(les_ctx_global, pushed(u, nb), pushed(v, nb), pushed(w, nb), &
         pushed(divtz, nb), dev(p, nbp), dev(dpdx, nb), dev(dpdy, nb)), 'press_stag_array')
    call pull(p, nbp)
    call pull(dpdx, nb)
    call pull(dpdy, nb)
  end subroutine les_h_press_stag_array

  ! forcing.f90:128-219 : project() with sim_param arrays explicit
  subroutine les_h_project(u, v, w, dpdx, dpdy, dpdz, nz)
    real(c_double), target :: u(*), v(*), w(*), dpdx(*), dpdy(*), dpdz(*)
    integer, intent(in) :: nz
    integer(c_int64_t) :: nb
    nb = field_bytes(nz)
    call les_check(les_project(les_ctx_global, pushed(u, nb), pushed(v, nb), pushed(w, nb), pushed(dpdx, nb), &
         pushed(dpdy, nb), pushed(dpdz, nb)), 'project')
    call pull(u, nb)
    call pull(v, nb)
    call pull(w, nb)
  end subroutine les_h_project

end module les_amd_host
