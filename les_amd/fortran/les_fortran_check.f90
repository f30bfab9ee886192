! Fortran-side check of the binding: the host-array wrappers of les_amd_host
! against closed forms -- Fourier modes through the spectral operators, a
! linear profile through ddz_uv, the 2 Delta cutoff of test_filter, the
! dealias1 / dealias2 round trip, a tridiagonal system with a known solution,
! the log law of wallstress on a uniform flow, Nu_t = |S| Co^2 l^2 with
! Mason's l for a uniform shear under model 1, the scale invariance of
! std_dynamic, Beta = 1 of lagrange_Ssim, calcbeta and step_scalar on uniform
! fields, rmsdiv of a divergence-free pair, timestep_conditions, the
! spectrum of one mode, the plane mean of avg_stats and the cosine sponge --
! then whole steps on the device state through les_step with every SGS model,
! the forcing and the convective step.  With a directory as its argument it
! writes each whole-step state there as a checkpoint record (model_<m>.out,
! convective.out) for tests/test_fortran_gpu.py to compare.
! Prints "LES_FORTRAN_OK" on success.
program les_fortran_check
  use iso_c_binding
  use les_amd
  use les_amd_host
  implicit none
  integer, parameter :: nx = 32, ny = 32, nz = 9, ld = 2 * (nx / 2 + 1), lh = nx / 2 + 1
  real(c_double), parameter :: pi = 3.1415926535897932384626433_c_double
  real(c_double), parameter :: dt = 2.0e-4_c_double, z0 = 1.0e-4_c_double, vonk = 0.41_c_double
  type(les_grid) :: g
  real(c_double) :: dx, dz, worst
  integer :: nchecks
  character(len=1024) :: outdir

  outdir = ''
  if (command_argument_count() >= 1) call get_command_argument(1, outdir)
  g = les_grid(nx, ny, nz, 2 * pi, 2 * pi, 1.0_c_double, dt, z0, 1.0_c_double, 1, 1, 5)
  dx = 2 * pi / nx
  dz = 1.0_c_double / (nz - 1)
  worst = 0
  nchecks = 0
  call les_init_global(g)
  call check_spectral()
  call check_vertical_and_filters()
  call check_tridag()
  call check_wall_and_model1()
  call check_dynamic_models()
  call check_diagnostics()
  call les_destroy(les_ctx_global)
  call check_scalars()
  call check_sponge()
  call check_whole_steps()
  write (*, '(a,i0,a,es10.3)') 'LES_FORTRAN_OK closed-form checks: ', nchecks, ' worst error: ', worst

contains

  subroutine expect(err, tol, what)
    real(c_double), intent(in) :: err, tol
    character(*), intent(in) :: what
    nchecks = nchecks + 1
    worst = max(worst, err)
    if (.not. (err <= tol)) then
      write (*, '(a,a,a,es12.4,a,es10.2)') 'FAILED ', what, ': error ', err, ' > ', tol
      error stop 2
    end if
  end subroutine expect

  real(c_double) function xc(i)
    integer, intent(in) :: i
    xc = (i - 1) * 2 * pi / nx
  end function xc

  real(c_double) function yc(j)
    integer, intent(in) :: j
    yc = (j - 1) * 2 * pi / ny
  end function yc

  subroutine check_spectral()
    real(c_double), allocatable, target :: f(:, :, :), fx(:, :, :), fy(:, :, :)
    real(c_double) :: err
    integer :: i, j, k
    allocate (f(ld, ny, nz), fx(ld, ny, nz), fy(ld, ny, nz))
    f = 0
    do k = 1, nz
      do j = 1, ny
        do i = 1, nx
          f(i, j, k) = sin(3 * xc(i) + 2 * yc(j)) + 0.1_c_double * k
        end do
      end do
    end do
    call les_h_filt_da(f, fx, fy, nz)
    err = 0
    do k = 1, nz
      do j = 1, ny
        do i = 1, nx
          err = max(err, abs(fx(i, j, k) - 3 * cos(3 * xc(i) + 2 * yc(j))), &
               abs(fy(i, j, k) - 2 * cos(3 * xc(i) + 2 * yc(j))))
        end do
      end do
    end do
    call expect(err, 1e-12_c_double, 'filt_da on sin(3x + 2y)')
    fx = 0
    fy = 0
    call les_h_ddxy(fx, fy, f, nz)
    err = 0
    do j = 1, ny
      do i = 1, nx
        err = max(err, abs(fx(i, j, 2) - 3 * cos(3 * xc(i) + 2 * yc(j))), &
             abs(fy(i, j, 2) - 2 * cos(3 * xc(i) + 2 * yc(j))))
      end do
    end do
    call expect(err, 1e-12_c_double, 'ddxy on sin(3x + 2y)')
    call les_h_ddx(fx, f, nz)
    call les_h_ddy(fy, f, nz)
    err = maxval(abs(fx(1:nx, :, 3) - 3 * cos(3 * spread([(xc(i), i = 1, nx)], 2, ny) &
         + 2 * spread([(yc(j), j = 1, ny)], 1, nx))))
    call expect(err, 1e-12_c_double, 'ddx on sin(3x + 2y)')
  end subroutine check_spectral

  subroutine check_vertical_and_filters()
    real(c_double), allocatable, target :: f(:, :, :), dfdz(:, :, :), big(:, :, :), back(:, :, :)
    real(c_double), allocatable, target :: G(:, :), plane(:, :)
    real(c_double) :: err
    integer :: i, j, k
    allocate (f(ld, ny, nz), dfdz(ld, ny, nz), G(lh, ny), plane(ld, ny))
    f = 0
    do k = 1, nz
      do j = 1, ny
        do i = 1, nx
          f(i, j, k) = 0.5_c_double * (k - 1) + sin(xc(i))
        end do
      end do
    end do
    dfdz = 0
    call les_h_ddz_uv(dfdz, f, nz)
    call expect(maxval(abs(dfdz(1:nx, :, 2:nz - 1) - 0.5_c_double / dz)), 1e-10_c_double, &
         'ddz_uv of a linear profile')
    call expect(maxval(abs(dfdz(:, :, nz))), 0.0_c_double, 'ddz_uv lid level')
    ! the 2 Delta sharp cutoff keeps k = 1 and removes k = 15 (cutoff 8)
    call les_h_test_filter_init(2.0_c_double, G)
    plane = 0
    do j = 1, ny
      do i = 1, nx
        plane(i, j) = sin(xc(i)) + sin(15 * xc(i))
      end do
    end do
    call les_h_test_filter(plane, G)
    err = 0
    do j = 1, ny
      do i = 1, nx
        err = max(err, abs(plane(i, j) - sin(xc(i))))
      end do
    end do
    call expect(err, 1e-13_c_double, 'test_filter 2 Delta cutoff')
    ! dealias1 then dealias2 returns a band-limited field
    allocate (big(2 * ((3 * nx / 2) / 2 + 1), 3 * ny / 2, nz), back(ld, ny, nz))
    f = 0
    do k = 1, nz
      do j = 1, ny
        do i = 1, nx
          f(i, j, k) = cos(2 * xc(i) + yc(j)) + 0.3_c_double * sin(5 * yc(j)) + k
        end do
      end do
    end do
    call les_h_dealias1(f, big, nz)
    back = 0
    call les_h_dealias2(back, big, nz)
    call expect(maxval(abs(back(1:nx, :, :) - f(1:nx, :, :))), 1e-12_c_double, 'dealias1 -> dealias2 round trip')
  end subroutine check_vertical_and_filters

  subroutine check_tridag()
    ! a u(j-1) + b u(j) + c u(j+1) = r for a known complex u; tridag_array
    ! leaves the mean mode and the Nyquist row and column alone
    integer, parameter :: n = nz + 1
    real(c_double), allocatable, target :: a(:, :, :), b(:, :, :), c(:, :, :), r(:, :, :, :), u(:, :, :, :)
    real(c_double), allocatable :: ut(:, :, :, :)
    real(c_double) :: err
    integer :: i, j, k, q
    allocate (a(lh, ny, n), b(lh, ny, n), c(lh, ny, n), r(2, lh, ny, n), u(2, lh, ny, n), ut(2, lh, ny, n))
    do k = 1, n
      do j = 1, ny
        do i = 1, lh
          a(i, j, k) = -1.0_c_double - 0.01_c_double * i
          b(i, j, k) = 4.0_c_double + 0.1_c_double * j
          c(i, j, k) = -1.0_c_double + 0.02_c_double * k
          ut(1, i, j, k) = cos(0.3_c_double * i + 0.7_c_double * j + k)
          ut(2, i, j, k) = sin(0.5_c_double * i - 0.2_c_double * j + 2 * k)
        end do
      end do
    end do
    do k = 1, n
      do q = 1, 2
        r(q, :, :, k) = b(:, :, k) * ut(q, :, :, k)
        if (k > 1) r(q, :, :, k) = r(q, :, :, k) + a(:, :, k) * ut(q, :, :, k - 1)
        if (k < n) r(q, :, :, k) = r(q, :, :, k) + c(:, :, k) * ut(q, :, :, k + 1)
      end do
    end do
    u = 0
    call les_h_tridag_array(a, b, c, r, u)
    err = 0
    do k = 1, n
      do j = 1, ny
        if (j == ny / 2 + 1) cycle
        do i = 1, lh - 1
          if (i == 1 .and. j == 1) cycle
          err = max(err, maxval(abs(u(:, i, j, k) - ut(:, i, j, k))))
        end do
      end do
    end do
    call expect(err, 1e-12_c_double, 'tridag_array known solution')
  end subroutine check_tridag

  subroutine check_wall_and_model1()
    real(c_double), parameter :: U = 4.0_c_double, gam = 0.7_c_double, Co = 0.16_c_double
    real(c_double), allocatable, target :: uu(:, :, :), vv(:, :, :), dudz(:, :, :), dvdz(:, :, :)
    real(c_double), allocatable, target :: txz(:, :, :), tyz(:, :, :), ps(:, :), ps0(:, :), phi(:, :), ust(:, :)
    real(c_double), allocatable, target :: gr(:, :, :, :), ww(:, :, :), tau(:, :, :, :)
    type(les_h_sgs) :: s
    real(c_double) :: us, delta, zz, l2, err
    integer :: k
    allocate (uu(ld, ny, nz), vv(ld, ny, nz), dudz(ld, ny, nz), dvdz(ld, ny, nz), txz(ld, ny, nz), &
         tyz(ld, ny, nz), ps(nx, ny), ps0(nx, ny), phi(nx, ny), ust(nx, ny))
    uu = U
    vv = 0; dudz = 0; dvdz = 0; txz = 0; tyz = 0; ps = 0; ps0 = 0; phi = 1
    call les_h_wallstress(uu, vv, dudz, dvdz, txz, tyz, ps, ps0, phi, ust, nz)
    us = U * vonk / log(0.5_c_double * dz / z0)
    call expect(maxval(abs(ust - us)) / us, 1e-14_c_double, 'wallstress u* of the log law')
    call expect(maxval(abs(txz(1:nx, :, 1) + us * us)) / (us * us), 1e-14_c_double, 'wallstress txz = -u*^2')
    call expect(maxval(abs(dudz(1:nx, :, 1) - us / (0.5_c_double * dz * vonk))) * dz, 1e-13_c_double, &
         'wallstress dudz')
    ! model 1 with a uniform shear dudz = gam: |S| = gam everywhere and
    ! Nu_t = gam Co^2 l(jz)^2 with Mason's l (sgs_stag.f90:84-94)
    call les_check(les_set_sgs_model(les_ctx_global, 1), 'set_sgs_model 1')
    allocate (gr(ld, ny, nz, 9), ww(ld, ny, nz), tau(ld, ny, nz, 6))
    gr = 0
    gr(:, :, :, 3) = gam
    ww = 0
    tau = 0
    call les_h_sgs_alloc(s)
    call les_check(les_set_time(les_ctx_global, 3, 3), 'set_time')
    call les_h_sgs_stag(gr(:, :, :, 1), gr(:, :, :, 2), gr(:, :, :, 3), gr(:, :, :, 4), gr(:, :, :, 5), &
         gr(:, :, :, 6), gr(:, :, :, 7), gr(:, :, :, 8), gr(:, :, :, 9), uu, vv, ww, tau(:, :, :, 1), &
         tau(:, :, :, 2), tau(:, :, :, 3), tau(:, :, :, 4), tau(:, :, :, 5), tau(:, :, :, 6), s, nz)
    delta = sqrt(dx * dx)
    err = 0
    do k = 1, nz - 1
      if (k == 1) then
        zz = 0.5_c_double * dz
      else
        zz = (k - 1) * dz
      end if
      l2 = 1 / (Co**2 / (vonk * zz)**2 + 1 / delta**2)
      err = max(err, maxval(abs(s%Nu_t(1:nx, :, k) - gam * Co**2 * l2)) / (gam * Co**2 * l2))
    end do
    call expect(err, 1e-13_c_double, 'model 1 Nu_t = |S| Co^2 l^2')
    call expect(maxval(abs(s%Cs_opt2 - Co**2)), 0.0_c_double, 'model 1 Cs_opt2 = Co^2')
    call les_check(les_set_sgs_model(les_ctx_global, 5), 'set_sgs_model 5')
  end subroutine check_wall_and_model1

  ! a band-limited random-phase field with an energy spectrum ~ k^(-5/3)
  ! (modes |k| <= 10) and its x and y derivatives on level k (0-based)
  subroutine turb_plane(k, seed, f, fx, fy)
    integer, intent(in) :: k, seed
    real(c_double), intent(out) :: f(nx, ny), fx(nx, ny), fy(nx, ny)
    real(c_double) :: kk, ph, a, arg
    integer :: mx, my, i, j
    f = 0; fx = 0; fy = 0
    do mx = 0, 10
      do my = -10, 10
        if (mx == 0 .and. my <= 0) cycle
        kk = sqrt(real(mx * mx + my * my, c_double))
        if (kk > 10) cycle
        ph = modulo(mx * 7 + my * 13 + seed * 5 + k * 3, 17) / 17.0_c_double * 2 * pi
        a = kk**(-5.0_c_double / 6) / kk
        do j = 1, ny
          do i = 1, nx
            arg = mx * xc(i) + my * yc(j) + ph
            f(i, j) = f(i, j) + a * sin(arg)
            fx(i, j) = fx(i, j) + a * mx * cos(arg)
            fy(i, j) = fy(i, j) + a * my * cos(arg)
          end do
        end do
      end do
    end do
  end subroutine turb_plane

  subroutine check_dynamic_models()
    real(c_double), allocatable, target :: S(:, :, :, :), u(:, :, :), v(:, :, :), w(:, :, :)
    real(c_double), allocatable, target :: S2(:, :, :, :), u2(:, :, :), v2(:, :, :), w2(:, :, :)
    real(c_double), target :: cs(nz), cs2(nz)
    real(c_double) :: uu(nx, ny), ux(nx, ny), uy(nx, ny), vv(nx, ny), vx(nx, ny), vy(nx, ny)
    type(les_h_sgs) :: sg
    integer :: k
    allocate (S(ld, ny, nz, 6), u(ld, ny, nz), v(ld, ny, nz), w(ld, ny, nz))
    S = 0; u = 0; v = 0; w = 0
    do k = 1, nz
      call turb_plane(k - 1, 1, uu, ux, uy)
      call turb_plane(k - 1, 2, vv, vx, vy)
      u(1:nx, :, k) = 5 + uu
      v(1:nx, :, k) = vv
      S(1:nx, :, k, 1) = ux
      S(1:nx, :, k, 4) = vy
      S(1:nx, :, k, 2) = 0.5_c_double * (uy + vx)
    end do
    call les_h_plane_dynamic(S(:, :, :, 1), S(:, :, :, 2), S(:, :, :, 3), S(:, :, :, 4), S(:, :, :, 5), &
         S(:, :, :, 6), u, v, w, cs, nz)
    S2 = 2.5_c_double * S
    u2 = 2.5_c_double * u
    v2 = 2.5_c_double * v
    w2 = 2.5_c_double * w
    call les_h_plane_dynamic(S2(:, :, :, 1), S2(:, :, :, 2), S2(:, :, :, 3), S2(:, :, :, 4), S2(:, :, :, 5), &
         S2(:, :, :, 6), u2, v2, w2, cs2, nz)
    call expect(maxval(abs(cs2 - cs)) / max(maxval(abs(cs)), 1e-30_c_double), 1e-12_c_double, &
         'std_dynamic scale invariance')
    call expect(-min(minval(cs), 0.0_c_double), 0.0_c_double, 'std_dynamic Cs >= 0')
    call expect(merge(0.0_c_double, 1.0_c_double, maxval(cs) > 0.005_c_double), 0.0_c_double, &
         'std_dynamic Cs > 0 on a k^(-5/3) field')
    ! lagrange_Ssim: Beta = 1 and its plane sums nx*ny
    call les_h_sgs_alloc(sg)
    sg%F_LM = 0.01_c_double
    sg%F_MM = 0.2_c_double
    call les_check(les_set_time(les_ctx_global, 5, 600), 'set_time')
    call les_h_lagrange(S(:, :, :, 1), S(:, :, :, 2), S(:, :, :, 3), S(:, :, :, 4), S(:, :, :, 5), &
         S(:, :, :, 6), u, v, w, sg, nz, .true.)
    call expect(maxval(abs(sg%Beta - 1)), 0.0_c_double, 'lagrange_Ssim Beta = 1')
    call expect(maxval(abs(sg%Beta_avg - nx * ny)), 0.0_c_double, 'lagrange_Ssim Beta_avg = nx ny')
    call expect(-min(minval(sg%Cs_opt2) - 1e-32_c_double, 0.0_c_double), 0.0_c_double, 'lagrange_Ssim Cs >= eps')
  end subroutine check_dynamic_models

  subroutine check_diagnostics()
    real(c_double), allocatable, target :: a(:, :, :), b(:, :, :), c(:, :, :), nu(:, :, :), spec(:, :), st(:)
    real(c_double), target :: rms
    real(c_double) :: out(4), ex(4)
    integer :: i, j, k, nl
    allocate (a(ld, ny, nz), b(ld, ny, nz), c(ld, ny, nz), nu(ld, ny, nz))
    a = 0; b = 0; c = 0
    do k = 1, nz
      do j = 1, ny
        do i = 1, nx
          a(i, j, k) = cos(xc(i) + yc(j))
          b(i, j, k) = -cos(xc(i) + yc(j))
        end do
      end do
    end do
    call les_h_rmsdiv(a, b, c, rms, nz)
    call expect(rms, 1e-14_c_double, 'rmsdiv of a divergence-free pair')
    a = 3.0_c_double; b = -4.0_c_double; c = 0.5_c_double; nu = 0.02_c_double
    call les_h_timestep_conditions(a, b, c, nu, out, nz)
    ex = [sqrt(9.0_c_double + 16 + 0.25_c_double) * dt / dx, 4 * dt / dx, 0.5_c_double * dt / dz, &
         dt * 0.02_c_double / dx**2]
    call expect(maxval(abs(out - ex) / ex), 1e-14_c_double, 'timestep_conditions of uniform fields')
    ! spectrum of 2 cos(5x): E(5) = 1, the rest 0; avg_stats plane mean
    a = 0
    do k = 1, nz
      do j = 1, ny
        do i = 1, nx
          a(i, j, k) = 2 * cos(5 * xc(i))
        end do
      end do
    end do
    allocate (spec(nx / 2, nz))
    call les_h_spectrum(a, nz, spec)
    call expect(abs(spec(6, 1) - 1) + sum(abs(spec(1:5, :))) + sum(abs(spec(7:, :))), 1e-13_c_double, &
         'spectrum of one mode')
    nl = nz - 1
    allocate (st(nl * (9 + nx / 2)))
    a(1:nx, :, :) = a(1:nx, :, :) + 7
    b = 0; c = 0
    call les_h_avg_stats(a, b, a, c, st, nz)
    call expect(maxval(abs(st(1:nl) - 7)), 1e-13_c_double, 'avg_stats plane mean')
  end subroutine check_diagnostics

  subroutine check_scalars()
    real(c_double), allocatable, target :: th(:, :, :), q(:, :, :), beta(:, :, :), R(:, :, :), Rf(:, :, :)
    type(les_scalar_params) :: sp
    call les_init_global(g)
    sp = les_scalar_params(1, 5, 0, 0, 0.55_c_double, 1000.0_c_double, 1.0_c_double, 400.0_c_double, &
         100.0_c_double, 9.81_c_double, 0.0_c_double, 0.0_c_double)
    call les_check(les_set_scalars(les_ctx_global, sp, c_null_ptr, c_null_ptr), 'set_scalars')
    allocate (th(ld, ny, nz), q(ld, ny, nz), beta(ld, ny, nz), R(ld, ny, nz), Rf(ld, ny, nz))
    th = 0.7_c_double
    q = 0.01_c_double
    beta = 1
    call les_h_calcbeta(th, q, beta, nz)
    call expect(maxval(abs(beta(1:nx, :, 1:nz - 1))), 1e-14_c_double, 'calcbeta of a uniform theta')
    R = 0.3_c_double
    Rf = 0.3_c_double
    call les_h_step_scalar(0, th, R, Rf, nz)
    call expect(maxval(abs(th(1:nx, :, 1:nz - 1) - (0.7_c_double + dt * 0.3_c_double))), 1e-15_c_double, &
         'step_scalar with RHS = RHS_f')
    call les_destroy(les_ctx_global)
  end subroutine check_scalars

  subroutine check_sponge()
    ! damping_method 2 on nz = 13, L_z = 3000 m: level 11 sits at the middle
    ! of the layer above 2250 m, sponge = 3.9 / 2
    type(les_grid) :: g13
    type(les_forcing) :: f
    real(c_double), target :: sp(13)
    g13 = les_grid(nx, ny, 13, 2 * pi, 2 * pi, 3.0_c_double, dt, z0, 1.0_c_double, 1, 1, 5)
    call les_init_global(g13)
    f = les_forcing(0, 0.0_c_double, 0.0_c_double, 0.45_c_double, 1000.0_c_double, 1, 2)
    call les_check(les_set_forcing(les_ctx_global, f), 'set_forcing')
    call les_check(les_get_sponge(les_ctx_global, c_loc(sp)), 'get_sponge')
    call expect(abs(sp(11) - 1.95_c_double) + maxval(abs(sp(1:9))), 1e-14_c_double, 'cosine sponge profile')
    call les_destroy(les_ctx_global)
  end subroutine check_sponge

  subroutine check_whole_steps()
    real(c_double), allocatable, target :: uh(:, :, :), wt(:, :)
    type(les_scalar_params) :: sp
    type(les_forcing) :: f
    type(c_ptr) :: du
    integer(c_int), target :: nlev, es
    real(c_double), target :: tref, qref
    real(c_double) :: mean
    integer :: model
    character(len=1100) :: name
    allocate (uh(ld, ny, nz))
    do model = 1, 6
      call les_init_global(g)
      call les_check(les_set_sgs_model(les_ctx_global, model), 'set_sgs_model')
      if (model == 1) then
        f = les_forcing(1, 10.0_c_double, 0.0_c_double, 0.45_c_double, 1000.0_c_double, 1, 1)
        call les_check(les_set_forcing(les_ctx_global, f), 'set_forcing')
      end if
      call les_check(les_state_alloc(les_ctx_global, 0), 'state_alloc')
      call les_check(les_state_init_ic(les_ctx_global, 42_c_int64_t, 0.1_c_double, 1.0_c_double, 0.4_c_double), &
           'ic')
      call les_check(les_step(les_ctx_global, 6), 'step')
      du = les_state_field(les_ctx_global, 'u'//c_null_char, c_loc(nlev), c_loc(es))
      call les_check(les_memcpy_d2h(les_ctx_global, c_loc(uh), du, int(8 * ld * ny * nz, c_int64_t)), 'd2h')
      mean = sum(uh(1:nx, 1:ny, 1:nz - 1)) / (nx * ny * (nz - 1))
      call expect(merge(0.0_c_double, 1.0_c_double, mean > 1 .and. mean < 100), 0.0_c_double, &
           'plausible mean u after 6 steps')
      if (len_trim(outdir) > 0) then
        write (name, '(a,a,i0,a)') trim(outdir), '/model_', model, '.out'
        call les_check(les_state_checkpoint(les_ctx_global, trim(name)//c_null_char), 'checkpoint')
      end if
      call les_destroy(les_ctx_global)
    end do
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
    du = les_state_field(les_ctx_global, 'theta'//c_null_char, c_loc(nlev), c_loc(es))
    call les_check(les_memcpy_d2h(les_ctx_global, c_loc(uh), du, int(8 * ld * ny * nz, c_int64_t)), 'd2h theta')
    call les_check(les_scalar_ref(les_ctx_global, c_loc(tref), c_loc(qref)), 'scalar_ref')
    mean = (sum(uh(1:nx, 1:ny, 1)) / (nx * ny) + tref) * 400.0_c_double
    call expect(merge(0.0_c_double, 1.0_c_double, mean > 282 .and. mean < 284), 0.0_c_double, &
         'plausible surface theta after 5 convective steps')
    if (len_trim(outdir) > 0) then
      name = trim(outdir)//'/convective.out'
      call les_check(les_state_checkpoint(les_ctx_global, trim(name)//c_null_char), 'checkpoint convective')
    end if
    call les_destroy(les_ctx_global)
  end subroutine check_whole_steps

end program les_fortran_check
