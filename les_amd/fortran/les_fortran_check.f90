! Fortran-side check of the binding: filt_da through the host wrapper on a
! single Fourier mode (derivatives known in closed form), then whole steps on
! the device state through les_step: 5 neutral steps (model 5), 6 with SGS
! model 3, and 5 convective steps with theta and q.  Prints "LES_FORTRAN_OK"
! on success.
program les_fortran_check
  use iso_c_binding
  use les_amd
  use les_amd_host
  implicit none
  integer, parameter :: nx = 32, ny = 32, nz = 9, ld = 2 * (nx / 2 + 1)
  real(c_double), parameter :: pi = 3.1415926535897932384626433_c_double
  type(les_grid) :: g
  real(c_double), allocatable, target :: f(:, :, :), dfdx(:, :, :), dfdy(:, :, :), uh(:, :, :)
  real(c_double) :: x, y, err, mean, mean3, tmean, tref, qref
  real(c_double), allocatable, target :: wt(:, :)
  type(les_scalar_params) :: sp
  integer :: i, j, k, nlev, es
  type(c_ptr) :: du

  g = les_grid(nx, ny, nz, 2 * pi, 2 * pi, 1.0_c_double, 2.0e-4_c_double, 1.0e-4_c_double, 1.0_c_double, 1, 1, 5)
  call les_init_global(g)
  allocate (f(ld, ny, nz), dfdx(ld, ny, nz), dfdy(ld, ny, nz))
  f = 0
  do k = 1, nz
    do j = 1, ny
      do i = 1, nx
        x = (i - 1) * 2 * pi / nx
        y = (j - 1) * 2 * pi / ny
        f(i, j, k) = sin(3 * x + 2 * y)
      end do
    end do
  end do
  call les_h_filt_da(f, dfdx, dfdy, nz)
  err = 0
  do k = 1, nz
    do j = 1, ny
      do i = 1, nx
        x = (i - 1) * 2 * pi / nx
        y = (j - 1) * 2 * pi / ny
        err = max(err, abs(dfdx(i, j, k) - 3 * cos(3 * x + 2 * y)), abs(dfdy(i, j, k) - 2 * cos(3 * x + 2 * y)))
      end do
    end do
  end do
  if (err > 1e-12_c_double) then
    write (*, *) 'filt_da derivative error', err
    error stop 2
  end if
  call les_check(les_state_alloc(les_ctx_global, 0), 'state_alloc')
  call les_check(les_state_init_ic(les_ctx_global, 42_c_int64_t, 0.1_c_double, 1.0_c_double, 0.4_c_double), 'ic')
  call les_check(les_step(les_ctx_global, 5), 'step')
  allocate (uh(ld, ny, nz))
  du = les_state_field(les_ctx_global, 'u'//c_null_char, nlev, es)
  call les_check(les_memcpy_d2h(les_ctx_global, c_loc(uh), du, int(8 * ld * ny * nz, c_int64_t)), 'd2h')
  mean = sum(uh(1:nx, 1:ny, 1:nz - 1)) / (nx * ny * (nz - 1))
  if (.not. (mean > 1.0_c_double .and. mean < 100.0_c_double)) then
    write (*, *) 'implausible mean u after 5 steps', mean
    error stop 3
  end if
  call les_destroy(les_ctx_global)
  ! SGS model 3 (scaledep_dynamic) through the whole step
  call les_init_global(g)
  call les_check(les_set_sgs_model(les_ctx_global, 3), 'set_sgs_model')
  call les_check(les_state_alloc(les_ctx_global, 0), 'state_alloc model 3')
  call les_check(les_state_init_ic(les_ctx_global, 42_c_int64_t, 0.1_c_double, 1.0_c_double, 0.4_c_double), 'ic')
  call les_check(les_step(les_ctx_global, 6), 'step model 3')
  du = les_state_field(les_ctx_global, 'u'//c_null_char, nlev, es)
  call les_check(les_memcpy_d2h(les_ctx_global, c_loc(uh), du, int(8 * ld * ny * nz, c_int64_t)), 'd2h')
  mean3 = sum(uh(1:nx, 1:ny, 1:nz - 1)) / (nx * ny * (nz - 1))
  if (.not. (mean3 > 1.0_c_double .and. mean3 < 100.0_c_double)) then
    write (*, *) 'implausible mean u after 6 model-3 steps', mean3
    error stop 4
  end if
  call les_destroy(les_ctx_global)
  ! convective step: theta and q with a uniform surface heat flux
  call les_init_global(g)
  allocate (wt(nx, ny))
  wt = 0.05_c_double
  sp = les_scalar_params(1, 7, 0, 0, 0.55_c_double, 1000.0_c_double, 1.0_c_double, 400.0_c_double, &
       100.0_c_double, 9.81_c_double, 0.01_c_double, -0.004_c_double)
  call les_check(les_set_scalars(les_ctx_global, sp, c_loc(wt), c_null_ptr), 'set_scalars')
  call les_check(les_state_alloc(les_ctx_global, 0), 'state_alloc scalars')
  call les_check(les_state_init_ic(les_ctx_global, 42_c_int64_t, 0.1_c_double, 1.0_c_double, 0.4_c_double), 'ic')
  call les_check(les_state_init_scalar_ic(les_ctx_global, 42_c_int64_t, 283.0_c_double, 0.8_c_double, &
       0.01_c_double, 0.1_c_double, 1.0_c_double), 'scalar ic')
  call les_check(les_step(les_ctx_global, 5), 'step convective')
  du = les_state_field(les_ctx_global, 'theta'//c_null_char, nlev, es)
  call les_check(les_memcpy_d2h(les_ctx_global, c_loc(uh), du, int(8 * ld * ny * nz, c_int64_t)), 'd2h theta')
  call les_check(les_scalar_ref(les_ctx_global, tref, qref), 'scalar_ref')
  tmean = (sum(uh(1:nx, 1:ny, 1)) / (nx * ny) + tref) * 400.0_c_double
  if (.not. (tmean > 282.0_c_double .and. tmean < 284.0_c_double)) then
    write (*, *) 'implausible surface theta after 5 convective steps', tmean
    error stop 5
  end if
  write (*, '(a,es12.4,a,f10.5,a,f10.5,a,f10.4)') 'LES_FORTRAN_OK filt_da err=', err, ' mean u after 5 steps=', &
       mean, ' model 3:', mean3, ' theta(1) K:', tmean
  call les_destroy(les_ctx_global)
end program les_fortran_check
