! les_amd_capi.f90 -- generated by gen_capi.py from include/les_amd.h; do not edit.
! iso_c_binding interfaces of the whole les_amd C ABI (one per function).
module les_amd_capi
  use iso_c_binding
  implicit none

  type, bind(C) :: les_grid
    integer(c_int) :: nx
    integer(c_int) :: ny
    integer(c_int) :: nz
    real(c_double) :: Lx
    real(c_double) :: Ly
    real(c_double) :: Lz
    real(c_double) :: dt
    real(c_double) :: z0
    real(c_double) :: force
    integer(c_int) :: inilag
    integer(c_int) :: initu
    integer(c_int) :: cs_count
  end type les_grid

  type, bind(C) :: les_sgs_fields
    type(c_ptr) :: F_LM
    type(c_ptr) :: F_MM
    type(c_ptr) :: F_QN
    type(c_ptr) :: F_NN
    type(c_ptr) :: Beta
    type(c_ptr) :: Cs_opt2
    type(c_ptr) :: Nu_t
    type(c_ptr) :: u_lag
    type(c_ptr) :: v_lag
    type(c_ptr) :: w_lag
    type(c_ptr) :: TnLMMM
    type(c_ptr) :: TnQNNN
    type(c_ptr) :: L11t
    type(c_ptr) :: L22t
    type(c_ptr) :: L33t
    type(c_ptr) :: Q11t
    type(c_ptr) :: Q22t
    type(c_ptr) :: Q33t
    type(c_ptr) :: Beta_avg
    type(c_ptr) :: Betaclip_avg
  end type les_sgs_fields

  type, bind(C) :: les_forcing
    integer(c_int) :: coriolis_forcing
    real(c_double) :: ug0
    real(c_double) :: vg0
    real(c_double) :: u_star
    real(c_double) :: z_i
    integer(c_int) :: ubc
    integer(c_int) :: damping_method
  end type les_forcing

  type, bind(C) :: les_scalar_params
    integer(c_int) :: enable
    integer(c_int) :: model
    integer(c_int) :: initsc
    integer(c_int) :: passive
    real(c_double) :: Pr
    real(c_double) :: z_i
    real(c_double) :: u_star
    real(c_double) :: T_scale
    real(c_double) :: q_scale
    real(c_double) :: g
    real(c_double) :: dTdz_top
    real(c_double) :: dqdz_top
  end type les_scalar_params

  type, bind(C) :: les_scalar_sgs_fields
    type(c_ptr) :: G_LM
    type(c_ptr) :: G_MM
    type(c_ptr) :: G_QN
    type(c_ptr) :: G_NN
    type(c_ptr) :: u_lag_s
    type(c_ptr) :: v_lag_s
    type(c_ptr) :: w_lag_s
    type(c_ptr) :: Pr_t
    type(c_ptr) :: Cs2_Pr
    type(c_ptr) :: phi_h
    type(c_ptr) :: psi_m
  end type les_scalar_sgs_fields

  type, bind(C) :: les_param_nml
    real(c_double) :: q_scale
    real(c_double) :: T_scale
    real(c_double) :: u_star
    real(c_double) :: ug0
    real(c_double) :: vg0
    real(c_double) :: inv_strength
    real(c_double) :: inv_strength_q
    real(c_double) :: dTdz_top
    real(c_double) :: dqdz_top
    real(c_double) :: T_init
    real(c_double) :: q_init
    real(c_double) :: dt_dim
    real(c_double) :: z_i
    real(c_double) :: dt
    integer(c_int) :: coriolis_forcing
    integer(c_int) :: nsteps
    integer(c_int) :: c_count
    integer(c_int) :: p_count
    integer(c_int) :: ubc
    integer(c_int) :: damping_method
    integer(c_int) :: initu
    integer(c_int) :: initsc
    integer(c_int) :: inilag
    integer(c_int) :: jt_total
  end type les_param_nml

  interface
    integer(c_int) function les_create(g, out) bind(C, name='les_create')
      import
      implicit none
      type(les_grid), intent(in) :: g
      type(c_ptr) :: out
    end function
    subroutine les_destroy(c) bind(C, name='les_destroy')
      import
      implicit none
      type(c_ptr), value :: c
    end subroutine
    type(c_ptr) function les_last_error(c) bind(C, name='les_last_error')
      import
      implicit none
      type(c_ptr), value :: c
    end function
    integer(c_int) function les_set_stream(c, hip_stream) bind(C, name='les_set_stream')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: hip_stream
    end function
    integer(c_int) function les_get_grid(c, out) bind(C, name='les_get_grid')
      import
      implicit none
      type(c_ptr), value :: c
      type(les_grid) :: out
    end function
    integer(c_int) function les_sync(c) bind(C, name='les_sync')
      import
      implicit none
      type(c_ptr), value :: c
    end function
    integer(c_int) function les_set_time(c, jt, jt_total) bind(C, name='les_set_time')
      import
      implicit none
      type(c_ptr), value :: c
      integer(c_int), value :: jt
      integer(c_int), value :: jt_total
    end function
    integer(c_int) function les_get_time(c, jt, jt_total) bind(C, name='les_get_time')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: jt
      type(c_ptr), value :: jt_total
    end function
    integer(c_int) function les_set_lag_init(c, flm_init, fqn_init) bind(C, name='les_set_lag_init')
      import
      implicit none
      type(c_ptr), value :: c
      integer(c_int), value :: flm_init
      integer(c_int), value :: fqn_init
    end function
    integer(c_int) function les_get_tables(c, kx, ky, k2, G_test, G_test_test) bind(C, &
        name='les_get_tables')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: kx
      type(c_ptr), value :: ky
      type(c_ptr), value :: k2
      type(c_ptr), value :: G_test
      type(c_ptr), value :: G_test_test
    end function
    integer(c_int) function les_filt_da(c, f, dfdx, dfdy, nlev) bind(C, name='les_filt_da')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: f
      type(c_ptr), value :: dfdx
      type(c_ptr), value :: dfdy
      integer(c_int), value :: nlev
    end function
    integer(c_int) function les_ddx(c, dfdx, f, nlev) bind(C, name='les_ddx')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: dfdx
      type(c_ptr), value :: f
      integer(c_int), value :: nlev
    end function
    integer(c_int) function les_ddy(c, dfdy, f, nlev) bind(C, name='les_ddy')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: dfdy
      type(c_ptr), value :: f
      integer(c_int), value :: nlev
    end function
    integer(c_int) function les_ddxy(c, dfdx, dfdy, f, nlev) bind(C, name='les_ddxy')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: dfdx
      type(c_ptr), value :: dfdy
      type(c_ptr), value :: f
      integer(c_int), value :: nlev
    end function
    integer(c_int) function les_test_filter_init(c, alpha, G) bind(C, name='les_test_filter_init')
      import
      implicit none
      type(c_ptr), value :: c
      real(c_double), value :: alpha
      type(c_ptr), value :: G
    end function
    integer(c_int) function les_test_filter(c, plane, G) bind(C, name='les_test_filter')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: plane
      type(c_ptr), value :: G
    end function
    integer(c_int) function les_padd(c, u_big, u) bind(C, name='les_padd')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: u_big
      type(c_ptr), value :: u
    end function
    integer(c_int) function les_unpadd(c, cc, cc_big) bind(C, name='les_unpadd')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: cc
      type(c_ptr), value :: cc_big
    end function
    integer(c_int) function les_dealias1(c, u, u_big, nlev) bind(C, name='les_dealias1')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: u
      type(c_ptr), value :: u_big
      integer(c_int), value :: nlev
    end function
    integer(c_int) function les_dealias2(c, u, u_big, nlev) bind(C, name='les_dealias2')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: u
      type(c_ptr), value :: u_big
      integer(c_int), value :: nlev
    end function
    integer(c_int) function les_ddz_uv(c, dfdz, f) bind(C, name='les_ddz_uv')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: dfdz
      type(c_ptr), value :: f
    end function
    integer(c_int) function les_ddz_w(c, dfdz, f) bind(C, name='les_ddz_w')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: dfdz
      type(c_ptr), value :: f
    end function
    integer(c_int) function les_convec(c, u, v, w, dudy, dudz, dvdx, dvdz, dwdx, dwdy, cx, cy, &
        cz) bind(C, name='les_convec')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: u
      type(c_ptr), value :: v
      type(c_ptr), value :: w
      type(c_ptr), value :: dudy
      type(c_ptr), value :: dudz
      type(c_ptr), value :: dvdx
      type(c_ptr), value :: dvdz
      type(c_ptr), value :: dwdx
      type(c_ptr), value :: dwdy
      type(c_ptr), value :: cx
      type(c_ptr), value :: cy
      type(c_ptr), value :: cz
    end function
    integer(c_int) function les_wallstress(c, u, v, dudz, dvdz, txz, tyz, psi_m, psi_m0, phi_m, &
        ustar_avg) bind(C, name='les_wallstress')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: u
      type(c_ptr), value :: v
      type(c_ptr), value :: dudz
      type(c_ptr), value :: dvdz
      type(c_ptr), value :: txz
      type(c_ptr), value :: tyz
      type(c_ptr), value :: psi_m
      type(c_ptr), value :: psi_m0
      type(c_ptr), value :: phi_m
      type(c_ptr), value :: ustar_avg
    end function
    integer(c_int) function les_calc_sij(c, grad, S) bind(C, name='les_calc_sij')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr) :: grad(*)
      type(c_ptr) :: S(*)
    end function
    integer(c_int) function les_sgs_stag(c, grad, u, v, w, tau, s) bind(C, name='les_sgs_stag')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr) :: grad(*)
      type(c_ptr), value :: u
      type(c_ptr), value :: v
      type(c_ptr), value :: w
      type(c_ptr) :: tau(*)
      type(les_sgs_fields), intent(in) :: s
    end function
    integer(c_int) function les_lagrange_sdep(c, S, u, v, w, s_) bind(C, name='les_lagrange_sdep')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr) :: S(*)
      type(c_ptr), value :: u
      type(c_ptr), value :: v
      type(c_ptr), value :: w
      type(les_sgs_fields), intent(in) :: s_
    end function
    integer(c_int) function les_interpolag_sdep(c, s) bind(C, name='les_interpolag_sdep')
      import
      implicit none
      type(c_ptr), value :: c
      type(les_sgs_fields), intent(in) :: s
    end function
    integer(c_int) function les_set_sgs_model(c, model) bind(C, name='les_set_sgs_model')
      import
      implicit none
      type(c_ptr), value :: c
      integer(c_int), value :: model
    end function
    integer(c_int) function les_std_dynamic(c, S, u, v, w, Cs_opt2) bind(C, name='les_std_dynamic')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr) :: S(*)
      type(c_ptr), value :: u
      type(c_ptr), value :: v
      type(c_ptr), value :: w
      type(c_ptr), value :: Cs_opt2
    end function
    integer(c_int) function les_new_scaledep_dynamic(c, S, u, v, w, Cs_opt2) bind(C, &
        name='les_new_scaledep_dynamic')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr) :: S(*)
      type(c_ptr), value :: u
      type(c_ptr), value :: v
      type(c_ptr), value :: w
      type(c_ptr), value :: Cs_opt2
    end function
    integer(c_int) function les_lagrange_ssim(c, S, u, v, w, s_) bind(C, name='les_lagrange_ssim')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr) :: S(*)
      type(c_ptr), value :: u
      type(c_ptr), value :: v
      type(c_ptr), value :: w
      type(les_sgs_fields), intent(in) :: s_
    end function
    integer(c_int) function les_interpolag_ssim(c, s) bind(C, name='les_interpolag_ssim')
      import
      implicit none
      type(c_ptr), value :: c
      type(les_sgs_fields), intent(in) :: s
    end function
    integer(c_int) function les_set_forcing(c, f) bind(C, name='les_set_forcing')
      import
      implicit none
      type(c_ptr), value :: c
      type(les_forcing), intent(in) :: f
    end function
    integer(c_int) function les_get_sponge(c, sponge) bind(C, name='les_get_sponge')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: sponge
    end function
    integer(c_int) function les_scaledep_dynamic(c, S, u, v, w, Cs_opt2, beta) bind(C, &
        name='les_scaledep_dynamic')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr) :: S(*)
      type(c_ptr), value :: u
      type(c_ptr), value :: v
      type(c_ptr), value :: w
      type(c_ptr), value :: Cs_opt2
      type(c_ptr), value :: beta
    end function
    integer(c_int) function les_set_scalars(c, p, wt_s, wq_s) bind(C, name='les_set_scalars')
      import
      implicit none
      type(c_ptr), value :: c
      type(les_scalar_params), intent(in) :: p
      type(c_ptr), value :: wt_s
      type(c_ptr), value :: wq_s
    end function
    integer(c_int) function les_set_scalar_lag_init(c, glm_init, gqn_init) bind(C, &
        name='les_set_scalar_lag_init')
      import
      implicit none
      type(c_ptr), value :: c
      integer(c_int), value :: glm_init
      integer(c_int), value :: gqn_init
    end function
    integer(c_int) function les_scalar_ref(c, theta_ref, q_ref) bind(C, name='les_scalar_ref')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: theta_ref
      type(c_ptr), value :: q_ref
    end function
    integer(c_int) function les_obukhov(c, u, v, theta, qmix, sgs_t3, sgs_q3, psi_m, psi_m0, &
        psi_h, psi_h0, phi_m, phi_h, ustar_avg) bind(C, name='les_obukhov')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: u
      type(c_ptr), value :: v
      type(c_ptr), value :: theta
      type(c_ptr), value :: qmix
      type(c_ptr), value :: sgs_t3
      type(c_ptr), value :: sgs_q3
      type(c_ptr), value :: psi_m
      type(c_ptr), value :: psi_m0
      type(c_ptr), value :: psi_h
      type(c_ptr), value :: psi_h0
      type(c_ptr), value :: phi_m
      type(c_ptr), value :: phi_h
      type(c_ptr), value :: ustar_avg
    end function
    integer(c_int) function les_calcbeta(c, theta, qmix, beta_scal) bind(C, name='les_calcbeta')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: theta
      type(c_ptr), value :: qmix
      type(c_ptr), value :: beta_scal
    end function
    integer(c_int) function les_scalar_rhs_calc(c, which, dsdx, dsdy, dsdz, u, v, w, Nu_t, Pr_, &
        ustar_avg, phi_h, RHS, sgs_vert1) bind(C, name='les_scalar_rhs_calc')
      import
      implicit none
      type(c_ptr), value :: c
      integer(c_int), value :: which
      type(c_ptr), value :: dsdx
      type(c_ptr), value :: dsdy
      type(c_ptr), value :: dsdz
      type(c_ptr), value :: u
      type(c_ptr), value :: v
      type(c_ptr), value :: w
      type(c_ptr), value :: Nu_t
      type(c_ptr), value :: Pr_
      type(c_ptr), value :: ustar_avg
      type(c_ptr), value :: phi_h
      type(c_ptr), value :: RHS
      type(c_ptr), value :: sgs_vert1
    end function
    integer(c_int) function les_step_scalar(c, which, scalar, RHS, RHS_f) bind(C, &
        name='les_step_scalar')
      import
      implicit none
      type(c_ptr), value :: c
      integer(c_int), value :: which
      type(c_ptr), value :: scalar
      type(c_ptr), value :: RHS
      type(c_ptr), value :: RHS_f
    end function
    integer(c_int) function les_interpolag_scalar_sdep(c, s) bind(C, &
        name='les_interpolag_scalar_sdep')
      import
      implicit none
      type(c_ptr), value :: c
      type(les_scalar_sgs_fields), intent(in) :: s
    end function
    integer(c_int) function les_lagrange_scalar_sdep(c, S, u, v, w, theta, m, s_) bind(C, &
        name='les_lagrange_scalar_sdep')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr) :: S(*)
      type(c_ptr), value :: u
      type(c_ptr), value :: v
      type(c_ptr), value :: w
      type(c_ptr), value :: theta
      type(les_sgs_fields), intent(in) :: m
      type(les_scalar_sgs_fields), intent(in) :: s_
    end function
    integer(c_int) function les_scalar_scaledep_dynamic(c, S, u, v, w, theta, Cs_opt2, s_) bind(C, &
        name='les_scalar_scaledep_dynamic')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr) :: S(*)
      type(c_ptr), value :: u
      type(c_ptr), value :: v
      type(c_ptr), value :: w
      type(c_ptr), value :: theta
      type(c_ptr), value :: Cs_opt2
      type(les_scalar_sgs_fields), intent(in) :: s_
    end function
    integer(c_int) function les_sgs_stag_scalar(c, grad, u, v, w, theta, tau, m, s) bind(C, &
        name='les_sgs_stag_scalar')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr) :: grad(*)
      type(c_ptr), value :: u
      type(c_ptr), value :: v
      type(c_ptr), value :: w
      type(c_ptr), value :: theta
      type(c_ptr) :: tau(*)
      type(les_sgs_fields), intent(in) :: m
      type(les_scalar_sgs_fields), intent(in) :: s
    end function
    integer(c_int) function les_divstress_uv(c, divt, tx, ty, tz) bind(C, name='les_divstress_uv')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: divt
      type(c_ptr), value :: tx
      type(c_ptr), value :: ty
      type(c_ptr), value :: tz
    end function
    integer(c_int) function les_divstress_w(c, divt, tx, ty, tz) bind(C, name='les_divstress_w')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: divt
      type(c_ptr), value :: tx
      type(c_ptr), value :: ty
      type(c_ptr), value :: tz
    end function
    integer(c_int) function les_press_stag_array(c, u, v, w, divtz, p, dpdx, dpdy) bind(C, &
        name='les_press_stag_array')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: u
      type(c_ptr), value :: v
      type(c_ptr), value :: w
      type(c_ptr), value :: divtz
      type(c_ptr), value :: p
      type(c_ptr), value :: dpdx
      type(c_ptr), value :: dpdy
    end function
    integer(c_int) function les_tridag_array(c, a, b, cc, r, u) bind(C, name='les_tridag_array')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: a
      type(c_ptr), value :: b
      type(c_ptr), value :: cc
      type(c_ptr), value :: r
      type(c_ptr), value :: u
    end function
    integer(c_int) function les_tridag_array_noalloc(c, a, b, cc, r, u) bind(C, &
        name='les_tridag_array_noalloc')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: a
      type(c_ptr), value :: b
      type(c_ptr), value :: cc
      type(c_ptr), value :: r
      type(c_ptr), value :: u
    end function
    integer(c_int) function les_project(c, u, v, w, dpdx, dpdy, dpdz) bind(C, name='les_project')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: u
      type(c_ptr), value :: v
      type(c_ptr), value :: w
      type(c_ptr), value :: dpdx
      type(c_ptr), value :: dpdy
      type(c_ptr), value :: dpdz
    end function
    integer(c_int) function les_rmsdiv(c, dudx, dvdy, dwdz, rms) bind(C, name='les_rmsdiv')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: dudx
      type(c_ptr), value :: dvdy
      type(c_ptr), value :: dwdz
      type(c_ptr), value :: rms
    end function
    integer(c_int) function les_timestep_conditions(c, u, v, w, Nu_t, out) bind(C, &
        name='les_timestep_conditions')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: u
      type(c_ptr), value :: v
      type(c_ptr), value :: w
      type(c_ptr), value :: Nu_t
      real(c_double) :: out(*)
    end function
    integer(c_int) function les_filt_da_f32(c, f, dfdx, dfdy, nlev) bind(C, name='les_filt_da_f32')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: f
      type(c_ptr), value :: dfdx
      type(c_ptr), value :: dfdy
      integer(c_int), value :: nlev
    end function
    integer(c_int) function les_ddx_f32(c, dfdx, f, nlev) bind(C, name='les_ddx_f32')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: dfdx
      type(c_ptr), value :: f
      integer(c_int), value :: nlev
    end function
    integer(c_int) function les_ddy_f32(c, dfdy, f, nlev) bind(C, name='les_ddy_f32')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: dfdy
      type(c_ptr), value :: f
      integer(c_int), value :: nlev
    end function
    integer(c_int) function les_test_filter_f32(c, plane, G) bind(C, name='les_test_filter_f32')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: plane
      type(c_ptr), value :: G
    end function
    integer(c_int) function les_dealias1_f32(c, u, u_big, nlev) bind(C, name='les_dealias1_f32')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: u
      type(c_ptr), value :: u_big
      integer(c_int), value :: nlev
    end function
    integer(c_int) function les_dealias2_f32(c, u, u_big, nlev) bind(C, name='les_dealias2_f32')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: u
      type(c_ptr), value :: u_big
      integer(c_int), value :: nlev
    end function
    integer(c_int) function les_convec_f32(c, u, v, w, dudy, dudz, dvdx, dvdz, dwdx, dwdy, cx, cy, &
        cz) bind(C, name='les_convec_f32')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: u
      type(c_ptr), value :: v
      type(c_ptr), value :: w
      type(c_ptr), value :: dudy
      type(c_ptr), value :: dudz
      type(c_ptr), value :: dvdx
      type(c_ptr), value :: dvdz
      type(c_ptr), value :: dwdx
      type(c_ptr), value :: dwdy
      type(c_ptr), value :: cx
      type(c_ptr), value :: cy
      type(c_ptr), value :: cz
    end function
    integer(c_int) function les_divstress_uv_f32(c, divt, tx, ty, tz) bind(C, &
        name='les_divstress_uv_f32')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: divt
      type(c_ptr), value :: tx
      type(c_ptr), value :: ty
      type(c_ptr), value :: tz
    end function
    integer(c_int) function les_divstress_w_f32(c, divt, tx, ty, tz) bind(C, &
        name='les_divstress_w_f32')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: divt
      type(c_ptr), value :: tx
      type(c_ptr), value :: ty
      type(c_ptr), value :: tz
    end function
    integer(c_int) function les_press_stag_array_f32(c, u, v, w, divtz, p, dpdx, dpdy) bind(C, &
        name='les_press_stag_array_f32')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: u
      type(c_ptr), value :: v
      type(c_ptr), value :: w
      type(c_ptr), value :: divtz
      type(c_ptr), value :: p
      type(c_ptr), value :: dpdx
      type(c_ptr), value :: dpdy
    end function
    integer(c_int) function les_state_alloc(c, precision) bind(C, name='les_state_alloc')
      import
      implicit none
      type(c_ptr), value :: c
      integer(c_int), value :: precision
    end function
    type(c_ptr) function les_state_field(c, name, nlev, elsize) bind(C, name='les_state_field')
      import
      implicit none
      type(c_ptr), value :: c
      character(kind=c_char) :: name(*)
      type(c_ptr), value :: nlev
      type(c_ptr), value :: elsize
    end function
    integer(c_int) function les_state_init_ic(c, seed, amp, ustar, kappa_ic) bind(C, &
        name='les_state_init_ic')
      import
      implicit none
      type(c_ptr), value :: c
      integer(c_int64_t), value :: seed
      real(c_double), value :: amp
      real(c_double), value :: ustar
      real(c_double), value :: kappa_ic
    end function
    integer(c_int) function les_state_init_scalar_ic(c, seed, T_init, z_inv, inv_strength, amp_K, &
        q_init) bind(C, name='les_state_init_scalar_ic')
      import
      implicit none
      type(c_ptr), value :: c
      integer(c_int64_t), value :: seed
      real(c_double), value :: T_init
      real(c_double), value :: z_inv
      real(c_double), value :: inv_strength
      real(c_double), value :: amp_K
      real(c_double), value :: q_init
    end function
    integer(c_int) function les_step(c, nsteps) bind(C, name='les_step')
      import
      implicit none
      type(c_ptr), value :: c
      integer(c_int), value :: nsteps
    end function
    integer(c_int) function les_set_graph(c, enable) bind(C, name='les_set_graph')
      import
      implicit none
      type(c_ptr), value :: c
      integer(c_int), value :: enable
    end function
    integer(c_int) function les_graph_prepare(c) bind(C, name='les_graph_prepare')
      import
      implicit none
      type(c_ptr), value :: c
    end function
    integer(c_int) function les_state_diag(c, rms, out4) bind(C, name='les_state_diag')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: rms
      real(c_double) :: out4(*)
    end function
    integer(c_int) function les_spectrum(c, f, nlev, spec) bind(C, name='les_spectrum')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: f
      integer(c_int), value :: nlev
      type(c_ptr), value :: spec
    end function
    integer(c_int) function les_avg_stats(c, u, v, w, txz, stats) bind(C, name='les_avg_stats')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: u
      type(c_ptr), value :: v
      type(c_ptr), value :: w
      type(c_ptr), value :: txz
      type(c_ptr), value :: stats
    end function
    integer(c_int) function les_surface_stats(c, u, v, txz, tyz, psi_m, psi_m0, ustar_avg, theta, &
        qmix, sgs_t3, sgs_q3, out8) bind(C, name='les_surface_stats')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: u
      type(c_ptr), value :: v
      type(c_ptr), value :: txz
      type(c_ptr), value :: tyz
      type(c_ptr), value :: psi_m
      type(c_ptr), value :: psi_m0
      type(c_ptr), value :: ustar_avg
      type(c_ptr), value :: theta
      type(c_ptr), value :: qmix
      type(c_ptr), value :: sgs_t3
      type(c_ptr), value :: sgs_q3
      type(c_ptr), value :: out8
    end function
    integer(c_int) function les_state_stats(c, stats, surface) bind(C, name='les_state_stats')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: stats
      type(c_ptr), value :: surface
    end function
    integer(c_int) function les_state_stats_accumulate(c) bind(C, name='les_state_stats_accumulate')
      import
      implicit none
      type(c_ptr), value :: c
    end function
    integer(c_int) function les_state_stats_average(c, stats, surface, n, reset) bind(C, &
        name='les_state_stats_average')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: stats
      type(c_ptr), value :: surface
      type(c_ptr), value :: n
      integer(c_int), value :: reset
    end function
    integer(c_int) function les_record_write(path, append, nseg, seg, bytes) bind(C, &
        name='les_record_write')
      import
      implicit none
      character(kind=c_char) :: path(*)
      integer(c_int), value :: append
      integer(c_int), value :: nseg
      type(c_ptr) :: seg(*)
      type(c_ptr), value :: bytes
    end function
    integer(c_int) function les_record_read(path, index, buf, buf_bytes, record_bytes) bind(C, &
        name='les_record_read')
      import
      implicit none
      character(kind=c_char) :: path(*)
      integer(c_int), value :: index
      type(c_ptr), value :: buf
      integer(c_int64_t), value :: buf_bytes
      type(c_ptr), value :: record_bytes
    end function
    integer(c_int) function les_state_checkpoint(c, path) bind(C, name='les_state_checkpoint')
      import
      implicit none
      type(c_ptr), value :: c
      character(kind=c_char) :: path(*)
    end function
    integer(c_int) function les_state_read_initial(c, path, model, initsc, inilag) bind(C, &
        name='les_state_read_initial')
      import
      implicit none
      type(c_ptr), value :: c
      character(kind=c_char) :: path(*)
      integer(c_int), value :: model
      integer(c_int), value :: initsc
      integer(c_int), value :: inilag
    end function
    integer(c_int) function les_read_total_time(path, jt_total) bind(C, name='les_read_total_time')
      import
      implicit none
      character(kind=c_char) :: path(*)
      type(c_ptr), value :: jt_total
    end function
    integer(c_int) function les_write_total_time(path, jt_total) bind(C, &
        name='les_write_total_time')
      import
      implicit none
      character(kind=c_char) :: path(*)
      integer(c_int), value :: jt_total
    end function
    integer(c_int) function les_read_surface(dir, nx, ny, z_i, wt_s, wq_s, zo) bind(C, &
        name='les_read_surface')
      import
      implicit none
      character(kind=c_char) :: dir(*)
      integer(c_int), value :: nx
      integer(c_int), value :: ny
      real(c_double), value :: z_i
      type(c_ptr), value :: wt_s
      type(c_ptr), value :: wq_s
      type(c_ptr), value :: zo
    end function
    integer(c_int) function les_read_param_nml(nml_path, total_time_path, p) bind(C, &
        name='les_read_param_nml')
      import
      implicit none
      character(kind=c_char) :: nml_path(*)
      character(kind=c_char) :: total_time_path(*)
      type(les_param_nml) :: p
    end function
    integer(c_int) function les_read_nml_value(path, group, key, value) bind(C, &
        name='les_read_nml_value')
      import
      implicit none
      character(kind=c_char) :: path(*)
      character(kind=c_char) :: group(*)
      character(kind=c_char) :: key(*)
      type(c_ptr), value :: value
    end function
    integer(c_int) function les_dev_alloc(c, bytes, ptr) bind(C, name='les_dev_alloc')
      import
      implicit none
      type(c_ptr), value :: c
      integer(c_int64_t), value :: bytes
      type(c_ptr) :: ptr
    end function
    integer(c_int) function les_dev_free(c, ptr) bind(C, name='les_dev_free')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: ptr
    end function
    integer(c_int) function les_memcpy_h2d(c, dst, src, bytes) bind(C, name='les_memcpy_h2d')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: dst
      type(c_ptr), value :: src
      integer(c_int64_t), value :: bytes
    end function
    integer(c_int) function les_memcpy_d2h(c, dst, src, bytes) bind(C, name='les_memcpy_d2h')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: dst
      type(c_ptr), value :: src
      integer(c_int64_t), value :: bytes
    end function
    type(c_ptr) function les_mirror(c, host, bytes) bind(C, name='les_mirror')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: host
      integer(c_int64_t), value :: bytes
    end function
    integer(c_int) function les_mirror_push(c, host, bytes) bind(C, name='les_mirror_push')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: host
      integer(c_int64_t), value :: bytes
    end function
    integer(c_int) function les_mirror_pull(c, host, bytes) bind(C, name='les_mirror_pull')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: host
      integer(c_int64_t), value :: bytes
    end function
    integer(c_int) function les_profile(c, enable) bind(C, name='les_profile')
      import
      implicit none
      type(c_ptr), value :: c
      integer(c_int), value :: enable
    end function
    integer(c_int) function les_profile_report(c, buf, buflen) bind(C, name='les_profile_report')
      import
      implicit none
      type(c_ptr), value :: c
      character(kind=c_char) :: buf(*)
      integer(c_int), value :: buflen
    end function
    integer(c_int) function les_autotune(c, what, report, report_len) bind(C, name='les_autotune')
      import
      implicit none
      type(c_ptr), value :: c
      integer(c_int), value :: what
      character(kind=c_char) :: report(*)
      integer(c_int), value :: report_len
    end function
    integer(c_int) function les_get_tune(c, out8) bind(C, name='les_get_tune')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: out8
    end function
    integer(c_int) function les_set_tune(c, in8) bind(C, name='les_set_tune')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: in8
    end function
    integer(c_int) function les_set_layout(c, xs_tiled) bind(C, name='les_set_layout')
      import
      implicit none
      type(c_ptr), value :: c
      integer(c_int), value :: xs_tiled
    end function
    integer(c_int) function les_get_layout(c) bind(C, name='les_get_layout')
      import
      implicit none
      type(c_ptr), value :: c
    end function
    integer(c_int) function les_ensemble_create(members, n, out) bind(C, name='les_ensemble_create')
      import
      implicit none
      type(c_ptr) :: members(*)
      integer(c_int), value :: n
      type(c_ptr) :: out
    end function
    subroutine les_ensemble_destroy(e) bind(C, name='les_ensemble_destroy')
      import
      implicit none
      type(c_ptr), value :: e
    end subroutine
    type(c_ptr) function les_ensemble_last_error(e) bind(C, name='les_ensemble_last_error')
      import
      implicit none
      type(c_ptr), value :: e
    end function
    integer(c_int) function les_ensemble_window(e) bind(C, name='les_ensemble_window')
      import
      implicit none
      type(c_ptr), value :: e
    end function
    integer(c_int) function les_ensemble_step(e, nsteps) bind(C, name='les_ensemble_step')
      import
      implicit none
      type(c_ptr), value :: e
      integer(c_int), value :: nsteps
    end function
    integer(c_int) function les_ensemble_sync(e) bind(C, name='les_ensemble_sync')
      import
      implicit none
      type(c_ptr), value :: e
    end function
    integer(c_int) function les_mlp_nparam() bind(C, name='les_mlp_nparam')
      import
      implicit none
    end function
    integer(c_int) function les_mlp_init(c, seed) bind(C, name='les_mlp_init')
      import
      implicit none
      type(c_ptr), value :: c
      integer(c_int64_t), value :: seed
    end function
    integer(c_int) function les_mlp_set_params(c, theta) bind(C, name='les_mlp_set_params')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: theta
    end function
    integer(c_int) function les_mlp_get_params(c, theta) bind(C, name='les_mlp_get_params')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: theta
    end function
    integer(c_int) function les_mlp_set_norm(c, mu, sd, ymu, ysd) bind(C, name='les_mlp_set_norm')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: mu
      type(c_ptr), value :: sd
      type(c_ptr), value :: ymu
      type(c_ptr), value :: ysd
    end function
    integer(c_int) function les_mlp_get_norm(c, mu, sd, ymu, ysd) bind(C, name='les_mlp_get_norm')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: mu
      type(c_ptr), value :: sd
      type(c_ptr), value :: ymu
      type(c_ptr), value :: ysd
    end function
    integer(c_int) function les_mlp_fit_norm(c, X, Y, P) bind(C, name='les_mlp_fit_norm')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: X
      type(c_ptr), value :: Y
      integer(c_int64_t), value :: P
    end function
    integer(c_int) function les_mlp_forward(c, X, P, Y) bind(C, name='les_mlp_forward')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: X
      integer(c_int64_t), value :: P
      type(c_ptr), value :: Y
    end function
    integer(c_int) function les_mlp_grad(c, X, T, P, grad, loss) bind(C, name='les_mlp_grad')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: X
      type(c_ptr), value :: T
      integer(c_int64_t), value :: P
      type(c_ptr), value :: grad
      type(c_ptr), value :: loss
    end function
    integer(c_int) function les_mlp_train(c, X, T, P, batch, steps, lr, seed, loss_hist) bind(C, &
        name='les_mlp_train')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: X
      type(c_ptr), value :: T
      integer(c_int64_t), value :: P
      integer(c_int), value :: batch
      integer(c_int), value :: steps
      real(c_float), value :: lr
      integer(c_int64_t), value :: seed
      type(c_ptr), value :: loss_hist
    end function
    integer(c_int) function les_mlp_harvest(c, X, Y) bind(C, name='les_mlp_harvest')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: X
      type(c_ptr), value :: Y
    end function
    integer(c_int) function les_set_closure(c, mode, period, threshold) bind(C, &
        name='les_set_closure')
      import
      implicit none
      type(c_ptr), value :: c
      integer(c_int), value :: mode
      integer(c_int), value :: period
      real(c_double), value :: threshold
    end function
    integer(c_int) function les_closure_judge(c, out) bind(C, name='les_closure_judge')
      import
      implicit none
      type(c_ptr), value :: c
      type(c_ptr), value :: out
    end function
  end interface
end module les_amd_capi
