! les_amd.f90 -- Fortran side of the C ABI in include/les_amd.h.
!
! module les_amd_capi : generated iso_c_binding interfaces of every C entry
!                        point and struct (les_amd_capi.f90, gen_capi.py).
! module les_amd      : les_amd_capi plus the global context and les_check.
! module les_amd_host : host-array wrappers with the reference routines'
!                        argument lists (filt_da.f90, ddx.f90, convec.f90,
!                        sgs_stag.f90 ...), one per routine-level entry point.
!                        Each keeps a device mirror per host array (les_mirror),
!                        pushes the inputs, runs the HIP routine and pulls the
!                        outputs, so a reference routine body can be replaced
!                        by one call.  Fields are (ld,ny,nlev) real(8), planes
!                        (ld,ny), surface arrays (nx,ny).  The whole-step path
!                        (les_step) keeps the state on the GPU instead and is
!                        the fast path.
module les_amd
  use iso_c_binding
  use les_amd_capi
  implicit none

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

  ! sgsmodule.f90 arrays of the Lagrangian models: (ld,ny,nz) fields and the
  ! per-level averages (nz)
  type :: les_h_sgs
    real(c_double), allocatable :: F_LM(:, :, :), F_MM(:, :, :), F_QN(:, :, :), F_NN(:, :, :), &
         Beta(:, :, :), Cs_opt2(:, :, :), Nu_t(:, :, :), u_lag(:, :, :), v_lag(:, :, :), w_lag(:, :, :), &
         TnLMMM(:, :, :), TnQNNN(:, :, :), L11t(:, :, :), L22t(:, :, :), L33t(:, :, :), Q11t(:, :, :), &
         Q22t(:, :, :), Q33t(:, :, :)
    real(c_double), allocatable :: Beta_avg(:), Betaclip_avg(:)
  end type les_h_sgs

  ! the scalar model's arrays: G_LM .. Pr_t, Cs2_Pr (ld,ny,nz); phi_h, psi_m (nx,ny)
  type :: les_h_scalar_sgs
    real(c_double), allocatable :: G_LM(:, :, :), G_MM(:, :, :), G_QN(:, :, :), G_NN(:, :, :), &
         u_lag_s(:, :, :), v_lag_s(:, :, :), w_lag_s(:, :, :), Pr_t(:, :, :), Cs2_Pr(:, :, :)
    real(c_double), allocatable :: phi_h(:, :), psi_m(:, :)
  end type les_h_scalar_sgs

contains

  ! ---------------------------------------------------------------- sizes
  type(les_grid) function grid()
    call les_check(les_get_grid(les_ctx_global, grid), 'les_get_grid')
  end function grid

  integer(c_int64_t) function field_bytes(nlev)
    integer, intent(in) :: nlev
    type(les_grid) :: g
    g = grid()
    field_bytes = int(2 * (g%nx / 2 + 1), c_int64_t) * g%ny * nlev * 8_c_int64_t
  end function field_bytes

  integer(c_int64_t) function surf_bytes()
    type(les_grid) :: g
    g = grid()
    surf_bytes = int(g%nx, c_int64_t) * g%ny * 8_c_int64_t
  end function surf_bytes

  integer(c_int64_t) function big_bytes(nlev)   ! (ld_big, ny2, nlev) of the 3/2 grid
    integer, intent(in) :: nlev
    type(les_grid) :: g
    g = grid()
    big_bytes = int(2 * ((3 * g%nx / 2) / 2 + 1), c_int64_t) * (3 * g%ny / 2) * nlev * 8_c_int64_t
  end function big_bytes

  integer(c_int64_t) function lhny_bytes()
    type(les_grid) :: g
    g = grid()
    lhny_bytes = int(g%nx / 2 + 1, c_int64_t) * g%ny * 8_c_int64_t
  end function lhny_bytes

  ! ---------------------------------------------------------------- mirrors
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

  ! ---------------------------------------------------------------- SGS state
  subroutine les_h_sgs_alloc(s)
    type(les_h_sgs), intent(inout) :: s
    type(les_grid) :: g
    integer :: ld
    g = grid()
    ld = 2 * (g%nx / 2 + 1)
    allocate (s%F_LM(ld, g%ny, g%nz), s%F_MM(ld, g%ny, g%nz), s%F_QN(ld, g%ny, g%nz), s%F_NN(ld, g%ny, g%nz), &
         s%Beta(ld, g%ny, g%nz), s%Cs_opt2(ld, g%ny, g%nz), s%Nu_t(ld, g%ny, g%nz), s%u_lag(ld, g%ny, g%nz), &
         s%v_lag(ld, g%ny, g%nz), s%w_lag(ld, g%ny, g%nz), s%TnLMMM(ld, g%ny, g%nz), s%TnQNNN(ld, g%ny, g%nz), &
         s%L11t(ld, g%ny, g%nz), s%L22t(ld, g%ny, g%nz), s%L33t(ld, g%ny, g%nz), s%Q11t(ld, g%ny, g%nz), &
         s%Q22t(ld, g%ny, g%nz), s%Q33t(ld, g%ny, g%nz), s%Beta_avg(g%nz), s%Betaclip_avg(g%nz))
    s%F_LM = 0; s%F_MM = 0; s%F_QN = 0; s%F_NN = 0; s%Beta = 0; s%Cs_opt2 = 0; s%Nu_t = 0
    s%u_lag = 0; s%v_lag = 0; s%w_lag = 0; s%TnLMMM = 0; s%TnQNNN = 0; s%L11t = 0; s%L22t = 0
    s%L33t = 0; s%Q11t = 0; s%Q22t = 0; s%Q33t = 0; s%Beta_avg = 0; s%Betaclip_avg = 0
  end subroutine les_h_sgs_alloc

  type(les_sgs_fields) function sgs_push(s)
    type(les_h_sgs), target, intent(in) :: s
    integer(c_int64_t) :: nb, nl
    nb = field_bytes(size(s%F_LM, 3))
    nl = 8_c_int64_t * size(s%Beta_avg)
    sgs_push = les_sgs_fields(pushed(s%F_LM, nb), pushed(s%F_MM, nb), pushed(s%F_QN, nb), pushed(s%F_NN, nb), &
         pushed(s%Beta, nb), pushed(s%Cs_opt2, nb), pushed(s%Nu_t, nb), pushed(s%u_lag, nb), pushed(s%v_lag, nb), &
         pushed(s%w_lag, nb), pushed(s%TnLMMM, nb), pushed(s%TnQNNN, nb), pushed(s%L11t, nb), pushed(s%L22t, nb), &
         pushed(s%L33t, nb), pushed(s%Q11t, nb), pushed(s%Q22t, nb), pushed(s%Q33t, nb), &
         pushed(s%Beta_avg, nl), pushed(s%Betaclip_avg, nl))
  end function sgs_push

  subroutine sgs_pull(s)
    type(les_h_sgs), target, intent(inout) :: s
    integer(c_int64_t) :: nb, nl
    nb = field_bytes(size(s%F_LM, 3))
    nl = 8_c_int64_t * size(s%Beta_avg)
    call pull(s%F_LM, nb); call pull(s%F_MM, nb); call pull(s%F_QN, nb); call pull(s%F_NN, nb)
    call pull(s%Beta, nb); call pull(s%Cs_opt2, nb); call pull(s%Nu_t, nb); call pull(s%u_lag, nb)
    call pull(s%v_lag, nb); call pull(s%w_lag, nb); call pull(s%TnLMMM, nb); call pull(s%TnQNNN, nb)
    call pull(s%L11t, nb); call pull(s%L22t, nb); call pull(s%L33t, nb); call pull(s%Q11t, nb)
    call pull(s%Q22t, nb); call pull(s%Q33t, nb); call pull(s%Beta_avg, nl); call pull(s%Betaclip_avg, nl)
  end subroutine sgs_pull

  subroutine les_h_scalar_sgs_alloc(s)
    type(les_h_scalar_sgs), intent(inout) :: s
    type(les_grid) :: g
    integer :: ld
    g = grid()
    ld = 2 * (g%nx / 2 + 1)
    allocate (s%G_LM(ld, g%ny, g%nz), s%G_MM(ld, g%ny, g%nz), s%G_QN(ld, g%ny, g%nz), s%G_NN(ld, g%ny, g%nz), &
         s%u_lag_s(ld, g%ny, g%nz), s%v_lag_s(ld, g%ny, g%nz), s%w_lag_s(ld, g%ny, g%nz), &
         s%Pr_t(ld, g%ny, g%nz), s%Cs2_Pr(ld, g%ny, g%nz), s%phi_h(g%nx, g%ny), s%psi_m(g%nx, g%ny))
    s%G_LM = 0; s%G_MM = 0; s%G_QN = 0; s%G_NN = 0; s%u_lag_s = 0; s%v_lag_s = 0; s%w_lag_s = 0
    s%Pr_t = 0; s%Cs2_Pr = 0; s%phi_h = 1; s%psi_m = 0
  end subroutine les_h_scalar_sgs_alloc

  type(les_scalar_sgs_fields) function ssgs_push(s)
    type(les_h_scalar_sgs), target, intent(in) :: s
    integer(c_int64_t) :: nb, n2
    nb = field_bytes(size(s%G_LM, 3))
    n2 = surf_bytes()
    ssgs_push = les_scalar_sgs_fields(pushed(s%G_LM, nb), pushed(s%G_MM, nb), pushed(s%G_QN, nb), &
         pushed(s%G_NN, nb), pushed(s%u_lag_s, nb), pushed(s%v_lag_s, nb), pushed(s%w_lag_s, nb), &
         pushed(s%Pr_t, nb), pushed(s%Cs2_Pr, nb), pushed(s%phi_h, n2), pushed(s%psi_m, n2))
  end function ssgs_push

  subroutine ssgs_pull(s)
    type(les_h_scalar_sgs), target, intent(inout) :: s
    integer(c_int64_t) :: nb
    nb = field_bytes(size(s%G_LM, 3))
    call pull(s%G_LM, nb); call pull(s%G_MM, nb); call pull(s%G_QN, nb); call pull(s%G_NN, nb)
    call pull(s%u_lag_s, nb); call pull(s%v_lag_s, nb); call pull(s%w_lag_s, nb); call pull(s%Pr_t, nb)
    call pull(s%Cs2_Pr, nb)
  end subroutine ssgs_pull

  ! ---------------------------------------------------------------- spectral
  ! filt_da.f90:3-42 : filt_da(f_c, dfdx_c, dfdy_c), f Nyquist-filtered in place
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

  ! ddxy.f90:1-29 : ddxy(dfdx, dfdy, f)
  subroutine les_h_ddxy(dfdx, dfdy, f, nlev)
    real(c_double), target :: dfdx(*), dfdy(*), f(*)
    integer, intent(in) :: nlev
    integer(c_int64_t) :: nb
    nb = field_bytes(nlev)
    call les_check(les_ddxy(les_ctx_global, dev(dfdx, nb), dev(dfdy, nb), pushed(f, nb), nlev), 'ddxy')
    call pull(dfdx, nb)
    call pull(dfdy, nb)
  end subroutine les_h_ddxy

  ! test_filtermodule.f90:32-63 : test_filter_init(alpha, G) with G (lh,ny)
  subroutine les_h_test_filter_init(alpha, G)
    real(c_double), intent(in) :: alpha
    real(c_double), target :: G(*)
    integer(c_int64_t) :: nb
    nb = lhny_bytes()
    call les_check(les_test_filter_init(les_ctx_global, alpha, dev(G, nb)), 'test_filter_init')
    call pull(G, nb)
  end subroutine les_h_test_filter_init

  ! test_filtermodule.f90:13-30 : test_filter(plane, G), (ld,ny) in place
  subroutine les_h_test_filter(plane, G)
    real(c_double), target :: plane(*), G(*)
    integer(c_int64_t) :: nb
    nb = field_bytes(1)
    call les_check(les_test_filter(les_ctx_global, pushed(plane, nb), pushed(G, lhny_bytes())), 'test_filter')
    call pull(plane, nb)
  end subroutine les_h_test_filter

  ! padd.f90:1-26 / unpadd.f90:1-22 : complex (lh,ny) <-> (lh_big,ny2) planes
  subroutine les_h_padd(u_big, u)
    real(c_double), target :: u_big(*), u(*)
    call les_check(les_padd(les_ctx_global, dev(u_big, big_bytes(1)), pushed(u, field_bytes(1))), 'padd')
    call pull(u_big, big_bytes(1))
  end subroutine les_h_padd

  subroutine les_h_unpadd(cc, cc_big)
    real(c_double), target :: cc(*), cc_big(*)
    call les_check(les_unpadd(les_ctx_global, dev(cc, field_bytes(1)), pushed(cc_big, big_bytes(1))), 'unpadd')
    call pull(cc, field_bytes(1))
  end subroutine les_h_unpadd

  ! dealias1.f90:1-31 (u untouched) / dealias2.f90:1-33
  subroutine les_h_dealias1(u, u_big, nlev)
    real(c_double), target :: u(*), u_big(*)
    integer, intent(in) :: nlev
    call les_check(les_dealias1(les_ctx_global, pushed(u, field_bytes(nlev)), dev(u_big, big_bytes(nlev)), nlev), &
         'dealias1')
    call pull(u_big, big_bytes(nlev))
  end subroutine les_h_dealias1

  subroutine les_h_dealias2(u, u_big, nlev)
    real(c_double), target :: u(*), u_big(*)
    integer, intent(in) :: nlev
    call les_check(les_dealias2(les_ctx_global, dev(u, field_bytes(nlev)), pushed(u_big, big_bytes(nlev)), nlev), &
         'dealias2')
    call pull(u, field_bytes(nlev))
    call pull(u_big, big_bytes(nlev))
  end subroutine les_h_dealias2

  ! ---------------------------------------------------------------- vertical FD
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

  ! ---------------------------------------------------------------- momentum
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

  ! wallstress.f90:5-69 ('wall') with the surface arrays explicit
  subroutine les_h_wallstress(u, v, dudz, dvdz, txz, tyz, psi_m, psi_m0, phi_m, ustar_avg, nz)
    real(c_double), target :: u(*), v(*), dudz(*), dvdz(*), txz(*), tyz(*)
    real(c_double), target :: psi_m(*), psi_m0(*), phi_m(*), ustar_avg(*)
    integer, intent(in) :: nz
    integer(c_int64_t) :: nb, n2
    nb = field_bytes(nz)
    n2 = surf_bytes()
    call les_check(les_wallstress(les_ctx_global, pushed(u, nb), pushed(v, nb), pushed(dudz, nb), &
         pushed(dvdz, nb), pushed(txz, nb), pushed(tyz, nb), pushed(psi_m, n2), pushed(psi_m0, n2), &
         pushed(phi_m, n2), dev(ustar_avg, n2)), 'wallstress')
    call pull(dudz, nb); call pull(dvdz, nb); call pull(txz, nb); call pull(tyz, nb); call pull(ustar_avg, n2)
  end subroutine les_h_wallstress

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

  ! tridag_array.f90:1-109 : a, b, c real (lh,ny,nz+1), r, u complex (lh,ny,nz+1)
  subroutine les_h_tridag_array(a, b, c, r, u)
    real(c_double), target :: a(*), b(*), c(*), r(*), u(*)
    type(les_grid) :: g
    integer(c_int64_t) :: nr
    g = grid()
    nr = int(g%nx / 2 + 1, c_int64_t) * g%ny * (g%nz + 1) * 8_c_int64_t
    call les_check(les_tridag_array(les_ctx_global, pushed(a, nr), pushed(b, nr), pushed(c, nr), &
         pushed(r, 2 * nr), dev(u, 2 * nr)), 'tridag_array')
    call pull(u, 2 * nr)
  end subroutine les_h_tridag_array

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

  ! rmsdiv.f90:1-113 : rmsdiv(rms) with the derivatives explicit
  subroutine les_h_rmsdiv(dudx, dvdy, dwdz, rms, nz)
    real(c_double), target :: dudx(*), dvdy(*), dwdz(*)
    real(c_double), target, intent(out) :: rms
    integer, intent(in) :: nz
    integer(c_int64_t) :: nb
    nb = field_bytes(nz)
    call les_check(les_rmsdiv(les_ctx_global, pushed(dudx, nb), pushed(dvdy, nb), pushed(dwdz, nb), c_loc(rms)), &
         'rmsdiv')
  end subroutine les_h_rmsdiv

  ! scalars_module2.f90:22-70 timestep_conditions: out = the four CFL and
  ! viscous stability numbers
  subroutine les_h_timestep_conditions(u, v, w, Nu_t, out, nz)
    real(c_double), target :: u(*), v(*), w(*), Nu_t(*)
    real(c_double), intent(out) :: out(4)
    integer, intent(in) :: nz
    integer(c_int64_t) :: nb
    nb = field_bytes(nz)
    call les_check(les_timestep_conditions(les_ctx_global, pushed(u, nb), pushed(v, nb), pushed(w, nb), &
         pushed(Nu_t, nb), out), 'timestep_conditions')
  end subroutine les_h_timestep_conditions

  ! ---------------------------------------------------------------- SGS
  ! sgs_stag.f90:44-53 calc_Sij(dudx, ..., dwdz) -> S11 .. S33
  subroutine les_h_calc_sij(dudx, dudy, dudz, dvdx, dvdy, dvdz, dwdx, dwdy, dwdz, &
       S11, S12, S13, S22, S23, S33, nz)
    real(c_double), target :: dudx(*), dudy(*), dudz(*), dvdx(*), dvdy(*), dvdz(*), dwdx(*), dwdy(*), dwdz(*)
    real(c_double), target :: S11(*), S12(*), S13(*), S22(*), S23(*), S33(*)
    integer, intent(in) :: nz
    integer(c_int64_t) :: nb
    type(c_ptr) :: gr(9), S(6)
    nb = field_bytes(nz)
    gr = [pushed(dudx, nb), pushed(dudy, nb), pushed(dudz, nb), pushed(dvdx, nb), pushed(dvdy, nb), &
         pushed(dvdz, nb), pushed(dwdx, nb), pushed(dwdy, nb), pushed(dwdz, nb)]
    S = [dev(S11, nb), dev(S12, nb), dev(S13, nb), dev(S22, nb), dev(S23, nb), dev(S33, nb)]
    call les_check(les_calc_sij(les_ctx_global, gr, S), 'calc_sij')
    call pull(S11, nb); call pull(S12, nb); call pull(S13, nb)
    call pull(S22, nb); call pull(S23, nb); call pull(S33, nb)
  end subroutine les_h_calc_sij

  ! sgs_stag.f90:3-366 with the selected model (les_set_sgs_model)
  subroutine les_h_sgs_stag(dudx, dudy, dudz, dvdx, dvdy, dvdz, dwdx, dwdy, dwdz, u, v, w, &
       txx, txy, txz, tyy, tyz, tzz, s, nz)
    real(c_double), target :: dudx(*), dudy(*), dudz(*), dvdx(*), dvdy(*), dvdz(*), dwdx(*), dwdy(*), dwdz(*)
    real(c_double), target :: u(*), v(*), w(*), txx(*), txy(*), txz(*), tyy(*), tyz(*), tzz(*)
    type(les_h_sgs), target, intent(inout) :: s
    integer, intent(in) :: nz
    integer(c_int64_t) :: nb
    type(c_ptr) :: gr(9), tau(6)
    nb = field_bytes(nz)
    gr = [pushed(dudx, nb), pushed(dudy, nb), pushed(dudz, nb), pushed(dvdx, nb), pushed(dvdy, nb), &
         pushed(dvdz, nb), pushed(dwdx, nb), pushed(dwdy, nb), pushed(dwdz, nb)]
    tau = [pushed(txx, nb), pushed(txy, nb), pushed(txz, nb), pushed(tyy, nb), pushed(tyz, nb), pushed(tzz, nb)]
    call les_check(les_sgs_stag(les_ctx_global, gr, pushed(u, nb), pushed(v, nb), pushed(w, nb), tau, &
         sgs_push(s)), 'sgs_stag')
    call pull(txx, nb); call pull(txy, nb); call pull(txz, nb)
    call pull(tyy, nb); call pull(tyz, nb); call pull(tzz, nb)
    call sgs_pull(s)
  end subroutine les_h_sgs_stag

  ! lagrange_Sdep.f90:4-478 / lagrange_Ssim.f90:4-300 (model 4, ssim = .true.)
  subroutine les_h_lagrange(S11, S12, S13, S22, S23, S33, u, v, w, s, nz, ssim)
    real(c_double), target :: S11(*), S12(*), S13(*), S22(*), S23(*), S33(*), u(*), v(*), w(*)
    type(les_h_sgs), target, intent(inout) :: s
    integer, intent(in) :: nz
    logical, intent(in) :: ssim
    integer(c_int64_t) :: nb
    type(c_ptr) :: Sp(6)
    nb = field_bytes(nz)
    Sp = [pushed(S11, nb), pushed(S12, nb), pushed(S13, nb), pushed(S22, nb), pushed(S23, nb), pushed(S33, nb)]
    if (ssim) then
      call les_check(les_lagrange_ssim(les_ctx_global, Sp, pushed(u, nb), pushed(v, nb), pushed(w, nb), &
           sgs_push(s)), 'lagrange_ssim')
    else
      call les_check(les_lagrange_sdep(les_ctx_global, Sp, pushed(u, nb), pushed(v, nb), pushed(w, nb), &
           sgs_push(s)), 'lagrange_sdep')
    end if
    call sgs_pull(s)
  end subroutine les_h_lagrange

  ! interpolag_Sdep.f90:2-380 / interpolag_Ssim.f90 (ssim = .true.)
  subroutine les_h_interpolag(s, ssim)
    type(les_h_sgs), target, intent(inout) :: s
    logical, intent(in) :: ssim
    if (ssim) then
      call les_check(les_interpolag_ssim(les_ctx_global, sgs_push(s)), 'interpolag_ssim')
    else
      call les_check(les_interpolag_sdep(les_ctx_global, sgs_push(s)), 'interpolag_sdep')
    end if
    call sgs_pull(s)
  end subroutine les_h_interpolag

  ! std_dynamic.f90:4-140 / scaledep_dynamic.f90:4-255 (beta present): the
  ! per-level Cs_opt2 (and beta) of the plane-averaged dynamic models
  subroutine les_h_plane_dynamic(S11, S12, S13, S22, S23, S33, u, v, w, cs, nz, beta)
    real(c_double), target :: S11(*), S12(*), S13(*), S22(*), S23(*), S33(*), u(*), v(*), w(*)
    real(c_double), target :: cs(*)
    integer, intent(in) :: nz
    real(c_double), target, optional :: beta(*)
    integer(c_int64_t) :: nb, nl
    type(c_ptr) :: S(6)
    nb = field_bytes(nz)
    nl = 8_c_int64_t * nz
    S = [pushed(S11, nb), pushed(S12, nb), pushed(S13, nb), pushed(S22, nb), pushed(S23, nb), pushed(S33, nb)]
    if (present(beta)) then
      call les_check(les_scaledep_dynamic(les_ctx_global, S, pushed(u, nb), pushed(v, nb), pushed(w, nb), &
           dev(cs, nl), dev(beta, nl)), 'scaledep_dynamic')
      call pull(beta, nl)
    else
      call les_check(les_std_dynamic(les_ctx_global, S, pushed(u, nb), pushed(v, nb), pushed(w, nb), dev(cs, nl)), &
           'std_dynamic')
    end if
    call pull(cs, nl)
  end subroutine les_h_plane_dynamic

  ! new_scaledep_dynamic.f90:1-223 (model 6): the per-level Cs_opt2
  subroutine les_h_new_scaledep_dynamic(S11, S12, S13, S22, S23, S33, u, v, w, cs, nz)
    real(c_double), target :: S11(*), S12(*), S13(*), S22(*), S23(*), S33(*), u(*), v(*), w(*)
    real(c_double), target :: cs(*)
    integer, intent(in) :: nz
    integer(c_int64_t) :: nb, nl
    type(c_ptr) :: S(6)
    nb = field_bytes(nz)
    nl = 8_c_int64_t * nz
    S = [pushed(S11, nb), pushed(S12, nb), pushed(S13, nb), pushed(S22, nb), pushed(S23, nb), pushed(S33, nb)]
    call les_check(les_new_scaledep_dynamic(les_ctx_global, S, pushed(u, nb), pushed(v, nb), pushed(w, nb), &
         dev(cs, nl)), 'new_scaledep_dynamic')
    call pull(cs, nl)
  end subroutine les_h_new_scaledep_dynamic

  ! ---------------------------------------------------------------- scalars
  ! obukhov, scalars_module.f90:508-690
  subroutine les_h_obukhov(u, v, theta, qmix, sgs_t3, sgs_q3, psi_m, psi_m0, psi_h, psi_h0, phi_m, phi_h, &
       ustar_avg, nz)
    real(c_double), target :: u(*), v(*), theta(*), qmix(*), sgs_t3(*), sgs_q3(*)
    real(c_double), target :: psi_m(*), psi_m0(*), psi_h(*), psi_h0(*), phi_m(*), phi_h(*), ustar_avg(*)
    integer, intent(in) :: nz
    integer(c_int64_t) :: nb, np, n2
    nb = field_bytes(nz)
    np = field_bytes(1)
    n2 = surf_bytes()
    call les_check(les_obukhov(les_ctx_global, pushed(u, nb), pushed(v, nb), pushed(theta, nb), pushed(qmix, nb), &
         pushed(sgs_t3, np), pushed(sgs_q3, np), pushed(psi_m, n2), pushed(psi_m0, n2), pushed(psi_h, n2), &
         pushed(psi_h0, n2), pushed(phi_m, n2), pushed(phi_h, n2), pushed(ustar_avg, n2)), 'obukhov')
    call pull(sgs_t3, np); call pull(sgs_q3, np); call pull(psi_m, n2); call pull(psi_m0, n2)
    call pull(psi_h, n2); call pull(psi_h0, n2); call pull(phi_m, n2); call pull(phi_h, n2)
    call pull(ustar_avg, n2)
  end subroutine les_h_obukhov

  ! calcbeta, scalars_module.f90:175-228
  subroutine les_h_calcbeta(theta, qmix, beta_scal, nz)
    real(c_double), target :: theta(*), qmix(*), beta_scal(*)
    integer, intent(in) :: nz
    integer(c_int64_t) :: nb
    nb = field_bytes(nz)
    call les_check(les_calcbeta(les_ctx_global, pushed(theta, nb), pushed(qmix, nb), dev(beta_scal, nb)), &
         'calcbeta')
    call pull(beta_scal, nb)
  end subroutine les_h_calcbeta

  ! scalar_RHS_calc, scalars_module.f90:232-454 (which = 0) and
  ! scalars_module_q.f90:114-330 (which = 1)
  subroutine les_h_scalar_rhs_calc(which, dsdx, dsdy, dsdz, u, v, w, Nu_t, Pr_, ustar_avg, phi_h, RHS, &
       sgs_vert1, nz)
    integer, intent(in) :: which, nz
    real(c_double), target :: dsdx(*), dsdy(*), dsdz(*), u(*), v(*), w(*), Nu_t(*), Pr_(*)
    real(c_double), target :: ustar_avg(*), phi_h(*), RHS(*), sgs_vert1(*)
    integer(c_int64_t) :: nb, n2
    nb = field_bytes(nz)
    n2 = surf_bytes()
    call les_check(les_scalar_rhs_calc(les_ctx_global, which, pushed(dsdx, nb), pushed(dsdy, nb), &
         pushed(dsdz, nb), pushed(u, nb), pushed(v, nb), pushed(w, nb), pushed(Nu_t, nb), pushed(Pr_, nb), &
         pushed(ustar_avg, n2), pushed(phi_h, n2), dev(RHS, nb), dev(sgs_vert1, field_bytes(1))), &
         'scalar_rhs_calc')
    call pull(dsdz, nb)
    call pull(RHS, nb)
    call pull(sgs_vert1, field_bytes(1))
  end subroutine les_h_scalar_rhs_calc

  ! step_scalar, scalars_module.f90:457-504 / step_scalar_q, scalars_module_q.f90:357-397
  subroutine les_h_step_scalar(which, scalar, RHS, RHS_f, nz)
    integer, intent(in) :: which, nz
    real(c_double), target :: scalar(*), RHS(*), RHS_f(*)
    integer(c_int64_t) :: nb
    nb = field_bytes(nz)
    call les_check(les_step_scalar(les_ctx_global, which, pushed(scalar, nb), pushed(RHS, nb), pushed(RHS_f, nb)), &
         'step_scalar')
    call pull(scalar, nb)
  end subroutine les_h_step_scalar

  ! interpolag_scalar_Sdep.f90:1-316
  subroutine les_h_interpolag_scalar_sdep(ss)
    type(les_h_scalar_sgs), target, intent(inout) :: ss
    call les_check(les_interpolag_scalar_sdep(les_ctx_global, ssgs_push(ss)), 'interpolag_scalar_sdep')
    call ssgs_pull(ss)
  end subroutine les_h_interpolag_scalar_sdep

  ! lagrange_scalar_Sdep.f90:1-467 (theta Nyquist-filtered in place)
  subroutine les_h_lagrange_scalar_sdep(S11, S12, S13, S22, S23, S33, u, v, w, theta, s, ss, nz)
    real(c_double), target :: S11(*), S12(*), S13(*), S22(*), S23(*), S33(*), u(*), v(*), w(*), theta(*)
    type(les_h_sgs), target, intent(inout) :: s
    type(les_h_scalar_sgs), target, intent(inout) :: ss
    integer, intent(in) :: nz
    integer(c_int64_t) :: nb
    type(c_ptr) :: Sp(6)
    nb = field_bytes(nz)
    Sp = [pushed(S11, nb), pushed(S12, nb), pushed(S13, nb), pushed(S22, nb), pushed(S23, nb), pushed(S33, nb)]
    call les_check(les_lagrange_scalar_sdep(les_ctx_global, Sp, pushed(u, nb), pushed(v, nb), pushed(w, nb), &
         pushed(theta, nb), sgs_push(s), ssgs_push(ss)), 'lagrange_scalar_sdep')
    call pull(theta, nb)
    call sgs_pull(s)
    call ssgs_pull(ss)
  end subroutine les_h_lagrange_scalar_sdep

  ! scalar_scaledep_dynamic.f90:1-307 (model 6): theta filtered in place,
  ! ss%Pr_t on levels 1..nz-1, the per-level Cs2_Pr in ss%Cs2_Pr(1:nz-1)
  subroutine les_h_scalar_scaledep_dynamic(S11, S12, S13, S22, S23, S33, u, v, w, theta, cs_opt2, ss, nz)
    real(c_double), target :: S11(*), S12(*), S13(*), S22(*), S23(*), S33(*), u(*), v(*), w(*)
    real(c_double), target :: theta(*), cs_opt2(*)
    type(les_h_scalar_sgs), target, intent(inout) :: ss
    integer, intent(in) :: nz
    integer(c_int64_t) :: nb
    type(c_ptr) :: S(6)
    nb = field_bytes(nz)
    S = [pushed(S11, nb), pushed(S12, nb), pushed(S13, nb), pushed(S22, nb), pushed(S23, nb), pushed(S33, nb)]
    call les_check(les_scalar_scaledep_dynamic(les_ctx_global, S, pushed(u, nb), pushed(v, nb), pushed(w, nb), &
         pushed(theta, nb), pushed(cs_opt2, nb), ssgs_push(ss)), 'scalar_scaledep_dynamic')
    call pull(theta, nb)
    call ssgs_pull(ss)
  end subroutine les_h_scalar_scaledep_dynamic

  ! sgs_stag.f90:3-366 with model = 7
  subroutine les_h_sgs_stag_scalar(dudx, dudy, dudz, dvdx, dvdy, dvdz, dwdx, dwdy, dwdz, u, v, w, theta, &
       txx, txy, txz, tyy, tyz, tzz, s, ss, nz)
    real(c_double), target :: dudx(*), dudy(*), dudz(*), dvdx(*), dvdy(*), dvdz(*), dwdx(*), dwdy(*), dwdz(*)
    real(c_double), target :: u(*), v(*), w(*), theta(*), txx(*), txy(*), txz(*), tyy(*), tyz(*), tzz(*)
    type(les_h_sgs), target, intent(inout) :: s
    type(les_h_scalar_sgs), target, intent(inout) :: ss
    integer, intent(in) :: nz
    integer(c_int64_t) :: nb
    type(c_ptr) :: gr(9), tau(6)
    nb = field_bytes(nz)
    gr = [pushed(dudx, nb), pushed(dudy, nb), pushed(dudz, nb), pushed(dvdx, nb), pushed(dvdy, nb), &
         pushed(dvdz, nb), pushed(dwdx, nb), pushed(dwdy, nb), pushed(dwdz, nb)]
    tau = [pushed(txx, nb), pushed(txy, nb), pushed(txz, nb), pushed(tyy, nb), pushed(tyz, nb), pushed(tzz, nb)]
    call les_check(les_sgs_stag_scalar(les_ctx_global, gr, pushed(u, nb), pushed(v, nb), pushed(w, nb), &
         pushed(theta, nb), tau, sgs_push(s), ssgs_push(ss)), 'sgs_stag_scalar')
    call pull(theta, nb)
    call pull(txx, nb); call pull(txy, nb); call pull(txz, nb)
    call pull(tyy, nb); call pull(tyz, nb); call pull(tzz, nb)
    call sgs_pull(s)
    call ssgs_pull(ss)
  end subroutine les_h_sgs_stag_scalar

  ! ---------------------------------------------------------------- statistics
  ! spectrum, io.f90:1613-1646 : spec (nx/2, nlev)
  subroutine les_h_spectrum(f, nlev, spec)
    real(c_double), target :: f(*), spec(*)
    integer, intent(in) :: nlev
    integer(c_int64_t) :: ns
    type(les_grid) :: g
    g = grid()
    ns = 8_c_int64_t * (g%nx / 2) * nlev
    call les_check(les_spectrum(les_ctx_global, pushed(f, field_bytes(nlev)), nlev, dev(spec, ns)), 'spectrum')
    call pull(spec, ns)
  end subroutine les_h_spectrum

  ! avg_stats, io.f90:1722-1755 : stats (9 nl + nl nx/2), nl = nz - 1
  subroutine les_h_avg_stats(u, v, w, txz, stats, nz)
    real(c_double), target :: u(*), v(*), w(*), txz(*), stats(*)
    integer, intent(in) :: nz
    integer(c_int64_t) :: nb, ns
    type(les_grid) :: g
    g = grid()
    nb = field_bytes(nz)
    ns = 8_c_int64_t * (nz - 1) * (9 + g%nx / 2)
    call les_check(les_avg_stats(les_ctx_global, pushed(u, nb), pushed(v, nb), pushed(w, nb), pushed(txz, nb), &
         dev(stats, ns)), 'avg_stats')
    call pull(stats, ns)
  end subroutine les_h_avg_stats

  ! obukhov's plane averages, scalars_module.f90:545-561,655-661 (neutral)
  subroutine les_h_surface_stats(u, v, txz, tyz, psi_m, psi_m0, ustar_avg, out8)
    real(c_double), target :: u(*), v(*), txz(*), tyz(*), psi_m(*), psi_m0(*), ustar_avg(*), out8(8)
    integer(c_int64_t) :: np, n2
    np = field_bytes(1)
    n2 = surf_bytes()
    call les_check(les_surface_stats(les_ctx_global, pushed(u, np), pushed(v, np), pushed(txz, np), &
         pushed(tyz, np), pushed(psi_m, n2), pushed(psi_m0, n2), pushed(ustar_avg, n2), c_null_ptr, c_null_ptr, &
         c_null_ptr, c_null_ptr, dev(out8, 64_c_int64_t)), 'surface_stats')
    call pull(out8, 64_c_int64_t)
  end subroutine les_h_surface_stats

end module les_amd_host
