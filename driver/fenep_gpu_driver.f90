! Fortran host driver of the MI355X FENE-P / IBM SMAC step.
!
! Replaces the step loop of the reference's `program main` (fenep.f90:2418-2569):
! the parameters, the grid and the output arrays are the reference's, the
! time step itself runs on the GPU through the C-ABI of include/fenep_gpu.h,
! bound here with iso_c_binding.  Outputs keep the reference's formats:
!   log line        log_progress          (fenep.f90:2266-2342)
!   koide3.d        trC at the centre     (fenep.f90:2529-2533)
!   Map/{z,y,x}_*.bin  18-column float64 planes (get_data_binary, fenep.f90:1108-1183)
!   Restart.d       U V W P C(1:6) Ax Ay Az Tx Ty Tz per cell (output/input,
!                   fenep.f90:1186-1233), written with 17 significant digits
!   Log/time_same.d 13-column time series of all_time_dif every Tstep steps
!                   (fenep_mpi.f90:2200-2348, read by _energy.py)
!   Log/energy_same_<step>.d  shell spectrum averaged over the Tstep records
!                   of each Gstep period (fft_energy_dif, fenep_mpi.f90:2566-2665)
!   Log/time_same_mask.d, Map/same_<step>.bin  masked time series and the
!                   18-column (x, y) map of all_time_dif_map (fenep_mpi.f90:2350-2564)
!   Log/energy_same_mask_<step>.d, Log/energy_same_z_<step>.d  masked z
!                   spectra of fft_energy_dif_map (fenep_mpi.f90:2670-2797)
!   the same records with 'diff' in place of 'same' on the second pass of
!                   input_type = 2, the time mean subtracted (fenep_mpi.f90:2999-3067)
!   ibm/<cylinder>.d  marker outline of each cylinder at k = NZ/2 (ibm_get,
!                   fenep.f90:1902-1917, read by ibm/ibm12/ibm3.py)
!   {z,y,all}_<step>.d  text maps, format (18e12.4) (get_data_xy, get_data_xz,
!                   get_data, fenep.f90:982-1106; _map.py form 'd'), map_text = 1
!
! Usage: fenep_gpu [namelist-file]      (default: fenep.nml)
module fenep_gpu_api
  use iso_c_binding
  implicit none

  type, bind(c) :: fp_params
    integer(c_int32_t) :: nx, ny, nz
    real(c_double) :: lx, ly, lz
    real(c_double) :: dt, re, beta, wi, lp
    integer(c_int32_t) :: method, walls_y, poisson_variant, precision
    integer(c_int32_t) :: newton_itrmax
    real(c_double) :: newton_eps, newton_dc
    integer(c_int32_t) :: sor_itrmax
    real(c_double) :: sor_eps, sor_omega
    real(c_double) :: body_force(3)
    integer(c_int32_t) :: n_dc_max
    real(c_double) :: dc_tol
    integer(c_int32_t) :: helm_sweeps, reserved0
    real(c_double) :: ibm_weight
    integer(c_int64_t) :: seed
    integer(c_int32_t) :: reserved(2)
  end type fp_params

  type, bind(c) :: fp_diag
    real(c_double) :: ke, cfl, trc_min, trc_max, trc_mean, trc_std, div_max, drag, lift
    real(c_double) :: newton_itr_mean, sor_itr, n_dc, dc_residual, sel_guarded, newton_unconverged
    integer(c_int32_t) :: nan_flag, step
  end type fp_diag

  interface
    integer(c_int) function fp_default_params(p) bind(c, name='fp_default_params')
      import :: c_int, fp_params
      type(fp_params), intent(out) :: p
    end function
    integer(c_int) function fp_create(p, ctx) bind(c, name='fp_create')
      import :: c_int, c_ptr, fp_params
      type(fp_params), intent(in) :: p
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function fp_time_stats(ctx, out) bind(c, name='fp_time_stats')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(13)
    end function
    integer(c_int) function fp_energy_spectrum(ctx, e, nbins) bind(c, name='fp_energy_spectrum')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: e(*)
      integer(c_int), value :: nbins
    end function
    integer(c_int) function fp_time_stats_map(ctx, ncyl, xy, diameter, map18, out13) &
        bind(c, name='fp_time_stats_map')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ncyl
      real(c_double), intent(in) :: xy(*)
      real(c_double), value :: diameter
      real(c_double), intent(out) :: map18(*), out13(13)
    end function
    integer(c_int) function fp_energy_spectrum_z(ctx, ncyl, xy, diameter, emask, efirst) &
        bind(c, name='fp_energy_spectrum_z')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ncyl
      real(c_double), intent(in) :: xy(*)
      real(c_double), value :: diameter
      real(c_double), intent(out) :: emask(*), efirst(*)
    end function
    integer(c_int) function fp_mean_accumulate(ctx) bind(c, name='fp_mean_accumulate')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function fp_mean_apply(ctx, mode) bind(c, name='fp_mean_apply')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode
    end function
    integer(c_int) function fp_destroy(ctx) bind(c, name='fp_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function fp_last_error(ctx) bind(c, name='fp_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function fp_set_state(ctx, u, v, w, p, c6, hist, has_history) bind(c, name='fp_set_state')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: u(*), v(*), w(*), p(*), c6(*)
      type(c_ptr), value :: hist
      integer(c_int), value :: has_history
    end function
    integer(c_int) function fp_get_state(ctx, u, v, w, p, c6, hist) bind(c, name='fp_get_state')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: u(*), v(*), w(*), p(*), c6(*)
      type(c_ptr), value :: hist
    end function
    integer(c_int) function fp_set_body_force(ctx, fx, fy, fz) bind(c, name='fp_set_body_force')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, fx, fy, fz
    end function
    integer(c_int) function fp_set_markers(ctx, nc, nl, xc, yc, zc, uc, vc, wc) bind(c, name='fp_set_markers')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nc, nl
      real(c_double), intent(in) :: xc(*), yc(*), zc(*), uc(*), vc(*), wc(*)
    end function
    integer(c_int) function fp_step(ctx, nsteps, sync) bind(c, name='fp_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps, sync
    end function
    integer(c_int) function fp_diagnostics(ctx, d) bind(c, name='fp_diagnostics')
      import :: c_int, c_ptr, fp_diag
      type(c_ptr), value :: ctx
      type(fp_diag), intent(out) :: d
    end function
    integer(c_int) function fp_seeded_uniform(seed, offset, n, out) bind(c, name='fp_seeded_uniform')
      import :: c_int, c_int64_t, c_size_t, c_double
      integer(c_int64_t), value :: seed, offset
      integer(c_size_t), value :: n
      real(c_double), intent(out) :: out(*)
    end function
  end interface

contains

  subroutine check(rc, ctx, what)
    integer(c_int), intent(in) :: rc
    type(c_ptr), intent(in) :: ctx
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: i
    if (rc == 0) return
    write(*, '(a, a, a, i0)', advance='no') 'fenep_gpu: ', what, ' failed, code ', rc
    if (c_associated(ctx)) then
      call c_f_pointer(fp_last_error(ctx), msg, [256])
      write(*, '(a)', advance='no') ': '
      do i = 1, 256
        if (msg(i) == c_null_char) exit
        write(*, '(a)', advance='no') msg(i)
      enddo
    endif
    write(*, *) ''
    stop 1
  end subroutine check

end module fenep_gpu_api


program fenep_gpu
  use iso_c_binding
  use fenep_gpu_api
  implicit none
  real(8), parameter :: PI = acos(-1.0d0)
  ! ---- namelist (reference parameter names, fenep.f90:4-49)
  integer :: NX = 32, NY = 32, NZ = 32
  real(8) :: Lx = 2*PI, Ly = 2*PI, Lz = 2*PI
  real(8) :: dt = 0.001d0, Re = 2.0d0, beta = 0.9d0, Wi = 2.0d0, Lp = 10.0d0
  integer :: method = 1, walls_y = 0, poisson_variant = 0, precision = 0
  integer :: Nstep = 100, Lstep = 10, Gstep = 100000000, Ostep = 100000000, koide_step = 20
  integer :: Tstep = 0
  integer :: flow_type = 0, input_type = 0
  integer(8) :: seed = 1
  real(8) :: force_amp = 0.1d0, body_force(3) = 0.0d0
  integer :: ncyl = 0
  real(8) :: DC = 0.5d0, xcyl(64) = 0.0d0, ycyl(64) = 0.0d0
  integer :: newton_itrmax = 10000, sor_itrmax = 10000, n_dc_max = 2
  real(8) :: newton_eps = 1.0d-6, sor_eps = 1.0d-5
  real(8) :: ibm_weight = 0.0d0      ! spreading weight: 0 = dX**3 (fenep), > 0 e.g. smac's dV
  integer :: map_text = 0
  character(256) :: dir = './'
  namelist /run/ NX, NY, NZ, Lx, Ly, Lz, dt, Re, beta, Wi, Lp, method, walls_y, poisson_variant, &
                 precision, Nstep, Lstep, Gstep, Ostep, koide_step, Tstep, flow_type, input_type, seed, &
                 force_amp, body_force, ncyl, DC, xcyl, ycyl, newton_itrmax, sor_itrmax, n_dc_max, &
                 newton_eps, sor_eps, ibm_weight, map_text, dir

  type(fp_params), target :: prm
  type(fp_diag) :: dg
  type(c_ptr) :: ctx = c_null_ptr
  real(8), allocatable, target :: U(:, :, :), V(:, :, :), W(:, :, :), P(:, :, :), C(:, :, :, :)
  real(8), allocatable, target :: hist(:, :, :, :), Fx(:, :, :), Fy(:, :, :), Fz(:, :, :)
  real(8), allocatable :: rnd(:), Xc(:), Yc(:), Zc(:), Uc(:), Vc(:), Wc(:)
  integer :: i, j, k, step, chunk, nl, rc, unit_nml
  character(256) :: nml_file
  real(8) :: dX, dY, dZ
  integer(8) :: t0, t1, trate
  integer(c_int) :: has_hist
  real(8), allocatable :: Energy(:), Etmp(:), Ezm(:), Ezf(:), Ezm_t(:), Ezf_t(:)
  integer :: n_energy = 0, n_energy_z = 0, pass
  character(4) :: key = 'same'
  real(c_double), allocatable :: cxy(:)

  nml_file = 'fenep.nml'
  if (command_argument_count() >= 1) call get_command_argument(1, nml_file)
  open(newunit=unit_nml, file=trim(nml_file), status='old', action='read')
  read(unit_nml, nml=run)
  close(unit_nml)
  if (dir(len_trim(dir):len_trim(dir)) /= '/') dir = trim(dir)//'/'
  call execute_command_line('mkdir -p '//trim(dir)//'Map')
  dX = Lx/NX; dY = Ly/NY; dZ = Lz/NZ

  rc = fp_default_params(prm)
  prm%nx = NX; prm%ny = NY; prm%nz = NZ
  prm%lx = Lx; prm%ly = Ly; prm%lz = Lz
  prm%dt = dt; prm%re = Re; prm%beta = beta; prm%wi = Wi; prm%lp = Lp
  prm%method = method; prm%walls_y = walls_y; prm%poisson_variant = poisson_variant
  prm%precision = precision
  prm%newton_itrmax = newton_itrmax; prm%newton_eps = newton_eps
  prm%sor_itrmax = sor_itrmax; prm%sor_eps = sor_eps; prm%n_dc_max = n_dc_max
  prm%body_force = body_force
  prm%seed = seed
  prm%ibm_weight = ibm_weight
  call check(fp_create(prm, ctx), ctx, 'fp_create')

  allocate(U(0:NX+1, 0:NY+1, 0:NZ+1), V(0:NX+1, 0:NY+1, 0:NZ+1), W(0:NX+1, 0:NY+1, 0:NZ+1))
  allocate(P(0:NX+1, 0:NY+1, 0:NZ+1), C(6, 0:NX+1, 0:NY+1, 0:NZ+1), hist(NX, NY, NZ, 9))
  allocate(Fx(NX, NY, NZ), Fy(NX, NY, NZ), Fz(NX, NY, NZ))

  ! ---- initial condition (init, fenep.f90:52-145) with a seeded stream
  call init_fields()
  has_hist = 0
  if (input_type > 0) then
    call read_restart()
    has_hist = 1
  endif
  if (has_hist == 1) then
    call check(fp_set_state(ctx, U, V, W, P, C, c_loc(hist), has_hist), ctx, 'fp_set_state')
  else
    call check(fp_set_state(ctx, U, V, W, P, C, c_null_ptr, 0), ctx, 'fp_set_state')
  endif
  if (flow_type == 3 .or. flow_type == 5) then
    call check(fp_set_body_force(ctx, c_loc(Fx), c_loc(Fy), c_loc(Fz)), ctx, 'fp_set_body_force')
  endif
  if (method == 2) call set_markers()

  write(*, '(a, 3i6, a, i2, a, i2)') 'grid:', NX, NY, NZ, '  method:', method, '  precision:', precision
  allocate(cxy(2*max(ncyl, 1)))
  cxy = 0.0d0
  do i = 1, ncyl
    cxy(2*i - 1) = xcyl(i); cxy(2*i) = ycyl(i)
  enddo
  if (method == 2) call ibm_get()
  call system_clock(t0, trate)
  ! input_type = 2 (fenep_mpi.f90:2999-3067): a first pass accumulates the
  ! time mean of the velocity, the second repeats the run from the same state
  ! with the statistics of the velocity less that mean (the 'diff' records)
  do pass = 1, merge(2, 1, input_type == 2)
    if (pass == 2) then
      call check(fp_mean_apply(ctx, 1_c_int), ctx, 'fp_mean_apply')
      key = 'diff'
      call init_fields()
      call read_restart()
      call check(fp_set_state(ctx, U, V, W, P, C, c_loc(hist), 1_c_int), ctx, 'fp_set_state')
    endif
    call run_pass()
  enddo
  call system_clock(t1)
  rc = fp_diagnostics(ctx, dg)
  if (dg%nan_flag /= 0) stop 'NaN value'
  write(*, '(a9, F10.3, a3)') 'Total   :', dble(t1 - t0)/dble(trate), '[s]'
  write(*, '(a, F12.2)') 'Mcell-steps/s:', dble(NX)*NY*NZ*Nstep/(dble(t1 - t0)/dble(trate))/1.0d6
  rc = fp_destroy(ctx)

contains

  subroutine run_pass()
    integer :: ev
    step = 0
    do while (step < Nstep)
      ev = next_event(step)
      if (input_type == 2 .and. pass == 1) ev = step + 1   ! the mean takes every step
      chunk = min(Nstep - step, ev - step)
      call check(fp_step(ctx, chunk, 1), ctx, 'fp_step')
      step = step + chunk
      if (input_type == 2 .and. pass == 1) call check(fp_mean_accumulate(ctx), ctx, 'fp_mean_accumulate')
      ! Fortran does not short-circuit .and.: test the period before mod()
      if (Lstep > 0) then
        if (mod(step, Lstep) == 0) call log_progress(step)
      endif
      if (koide_step > 0 .and. pass == 1) then
        if (mod(step, koide_step) == 0) call write_koide(step)
      endif
      if (Gstep > 0 .and. pass == 1) then
        if (mod(step, Gstep) == 0) then
          call get_data_binary(step)
          if (map_text > 0) call get_data_text(step)
        endif
      endif
      if (Ostep > 0 .and. pass == 1) then
        if (mod(step, Ostep) == 0) call write_restart()
      endif
      if (Tstep > 0) then
        if (mod(step, Tstep) == 0) then
          call write_time_stats(step)
          call write_time_stats_map(step)
          if (walls_y == 0) call add_energy()
          call add_energy_z()
        endif
      endif
      if (Gstep > 0 .and. Tstep > 0) then
        if (mod(step, Gstep) == 0 .and. n_energy > 0) call write_energy(step)
        if (mod(step, Gstep) == 0 .and. n_energy_z > 0) call write_energy_z(step)
      endif
    enddo
  end subroutine run_pass

  subroutine write_time_stats_map(s)
    ! all_time_dif_map (fenep_mpi.f90:2350-2564): Log/time_<key>_mask.d every
    ! record, Map/<key>_<step>.bin every Gstep
    integer, intent(in) :: s
    integer :: iu, r
    real(c_double) :: ts(13)
    real(c_double), allocatable :: mp(:)
    character(8) :: str
    allocate(mp(18*NX*NY))
    call check(fp_time_stats_map(ctx, int(ncyl, c_int), cxy, DC, mp, ts), ctx, 'fp_time_stats_map')
    call execute_command_line('mkdir -p '//trim(dir)//'Log')
    open(newunit=iu, file=trim(dir)//'Log/time_'//trim(key)//'_mask.d', position='append')
    write(iu, '(i10, 12es25.16e3)') s, s*dt, ts(3:13)
    close(iu)
    if (Gstep > 0) then
      if (mod(s, Gstep) == 0) then
        write(str, '(I8.8)') s
        open(newunit=iu, file=trim(dir)//'Map/'//trim(key)//'_'//str//'.bin', form='unformatted', &
             status='replace', access='stream')
        do r = 0, NX*NY - 1
          write(iu) mp(18*r + 1:18*r + 18)
        enddo
        close(iu)
      endif
    endif
    deallocate(mp)
  end subroutine write_time_stats_map

  subroutine add_energy_z()
    ! column z spectra of fft_energy_dif_map, accumulated between Gstep writes
    if (.not. allocated(Ezm)) then
      allocate(Ezm(0:NZ-1), Ezf(0:NZ-1), Ezm_t(0:NZ-1), Ezf_t(0:NZ-1))
      Ezm = 0.0d0; Ezf = 0.0d0
    endif
    call check(fp_energy_spectrum_z(ctx, int(ncyl, c_int), cxy, DC, Ezm_t, Ezf_t), ctx, 'fp_energy_spectrum_z')
    Ezm = Ezm + Ezm_t
    Ezf = Ezf + Ezf_t
    n_energy_z = n_energy_z + 1
  end subroutine add_energy_z

  subroutine write_energy_z(s)
    ! Log/energy_<key>_mask_<step>.d and Log/energy_<key>_z_<step>.d, rows
    ! kz = 0..NZ-1, averaged over the accumulated records (fenep_mpi.f90:2768-2794)
    integer, intent(in) :: s
    integer :: iu, k
    character(8) :: str
    write(str, '(I8.8)') s
    call execute_command_line('mkdir -p '//trim(dir)//'Log')
    open(newunit=iu, file=trim(dir)//'Log/energy_'//trim(key)//'_z_'//str//'.d', status='replace')
    do k = 0, NZ - 1
      write(iu, '(i8, es25.16e3)') k, Ezf(k)/dble(n_energy_z)
    enddo
    close(iu)
    open(newunit=iu, file=trim(dir)//'Log/energy_'//trim(key)//'_mask_'//str//'.d', status='replace')
    do k = 0, NZ - 1
      write(iu, '(i8, es25.16e3)') k, Ezm(k)/dble(n_energy_z)
    enddo
    close(iu)
    Ezm = 0.0d0; Ezf = 0.0d0
    n_energy_z = 0
  end subroutine write_energy_z

  subroutine ibm_get()
    ! the markers of each cylinder in the plane k = NZ/2, closed by the first
    ! marker (fenep.f90:1902-1917)
    integer :: iu, l, m, idx, nc
    character(8) :: str
    nc = max(ncyl, 1)
    call execute_command_line('mkdir -p '//trim(dir)//'ibm')
    do l = 1, nc
      write(str, '(I8.8)') l
      open(newunit=iu, file=trim(dir)//'ibm/'//str//'.d', status='replace')
      do m = 1, nl
        idx = l + nc*((m - 1) + nl*(NZ/2 - 1))
        write(iu, '(2e12.4)') Xc(idx), Yc(idx)
      enddo
      idx = l + nc*(nl*(NZ/2 - 1))
      write(iu, '(2e12.4)') Xc(idx), Yc(idx)
      close(iu)
    enddo
  end subroutine ibm_get

  integer function next_event(s)
    integer, intent(in) :: s
    integer :: e
    e = Nstep
    if (Lstep > 0) e = min(e, (s/Lstep + 1)*Lstep)
    if (koide_step > 0) e = min(e, (s/koide_step + 1)*koide_step)
    if (Gstep > 0) e = min(e, (s/Gstep + 1)*Gstep)
    if (Ostep > 0) e = min(e, (s/Ostep + 1)*Ostep)
    if (Tstep > 0) e = min(e, (s/Tstep + 1)*Tstep)
    next_event = e
  end function next_event

  subroutine uniform(offset, n, out)
    integer(8), intent(in) :: offset
    integer, intent(in) :: n
    real(8), intent(out) :: out(n)
    integer(c_int) :: r
    r = fp_seeded_uniform(int(seed, c_int64_t), int(offset, c_int64_t), int(n, c_size_t), out)
  end subroutine uniform

  subroutine init_fields()
    integer(8) :: nn
    nn = int(NX+2, 8)*(NY+2)*(NZ+2)
    allocate(rnd(6*nn))
    call uniform(0_8, int(nn), rnd(1:nn));            U = reshape(0.001d0*(rnd(1:nn) - 0.5d0), shape(U))
    call uniform(nn, int(nn), rnd(1:nn));             V = reshape(0.001d0*(rnd(1:nn) - 0.5d0), shape(V))
    call uniform(2*nn, int(nn), rnd(1:nn));           W = reshape(0.001d0*(rnd(1:nn) - 0.5d0), shape(W))
    call uniform(3*nn, int(6*nn), rnd);               C = reshape(0.001d0*(rnd - 0.5d0), shape(C))
    C(1, :, :, :) = C(1, :, :, :) + 1.0d0
    C(4, :, :, :) = C(4, :, :, :) + 1.0d0
    C(6, :, :, :) = C(6, :, :, :) + 1.0d0
    P = 0.0d0; Fx = 0.0d0; Fy = 0.0d0; Fz = 0.0d0; hist = 0.0d0
    if (walls_y /= 0) U = U + 1.0d0
    if (flow_type == 3 .or. flow_type == 4) then   ! Taylor-Green (fenep.f90:104-137)
      do k = 1, NZ
        do j = 1, NY
          do i = 1, NX
            if (flow_type == 3) then
              Fx(i, j, k) = -sin(i*dX)*cos((j - 0.5d0)*dY)
              Fy(i, j, k) = cos((i - 0.5d0)*dX)*sin(j*dY)
            endif
            U(i, j, k) = -sin(i*dX)*cos((j - 0.5d0)*dY)/(2.0d0/Re)
            V(i, j, k) = cos((i - 0.5d0)*dX)*sin(j*dY)/(2.0d0/Re)
          enddo
        enddo
      enddo
      if (flow_type == 3) then
        W = 0.0d0
        C = 0.0d0; C(1, :, :, :) = 1.0d0; C(4, :, :, :) = 1.0d0; C(6, :, :, :) = 1.0d0
      endif
    endif
    if (flow_type == 5) then   ! Kolmogorov forcing f_x = A sin(y) at the U points
      do k = 1, NZ
        do j = 1, NY
          Fx(:, j, k) = force_amp*sin((j - 0.5d0)*dY)
        enddo
      enddo
    endif
    deallocate(rnd)
  end subroutine init_fields

  subroutine set_markers()
    ! interior-shell markers of ibm_init (fenep.f90:1791-1801, 1857-1895) for
    ! stationary cylinders at (xcyl, ycyl)
    integer :: nlap, s, m, l, n, idx, nc
    integer, allocatable :: shell(:)
    real(8) :: rad
    nc = max(ncyl, 1)
    nlap = int(DC/(2*dX))
    allocate(shell(0:nlap))
    do s = 0, nlap
      shell(s) = nint(PI*(DC - 2*s*dX)/dX)
    enddo
    if (shell(nlap) == 0) shell(nlap) = 1
    nl = sum(shell)
    allocate(Xc(nc*nl*NZ), Yc(nc*nl*NZ), Zc(nc*nl*NZ), Uc(nc*nl*NZ), Vc(nc*nl*NZ), Wc(nc*nl*NZ))
    Uc = 0.0d0; Vc = 0.0d0; Wc = 0.0d0
    do n = 1, NZ
      m = 0
      do s = 0, nlap
        rad = (DC - 2*s*dX)/2
        do j = 1, shell(s)
          m = m + 1
          do l = 1, nc
            idx = l + nc*((m - 1) + nl*(n - 1))
            Xc(idx) = xcyl(l) + rad*cos(2*PI/shell(s)*j)
            Yc(idx) = ycyl(l) + rad*sin(2*PI/shell(s)*j)
            Zc(idx) = (n - 0.25d0)*dZ
          enddo
        enddo
      enddo
    enddo
    call check(fp_set_markers(ctx, int(nc, c_int), int(nl, c_int), Xc, Yc, Zc, Uc, Vc, Wc), ctx, 'fp_set_markers')
  end subroutine set_markers

  subroutine log_progress(s)
    integer, intent(in) :: s
    rc = fp_diagnostics(ctx, dg)
    write(*, '(a, I8)', advance='no') 'step:', s
    write(*, '(a, e12.4)', advance='no') '  | K_energy:', dg%ke
    write(*, '(a, F7.3)', advance='no') '  | CFL:', dg%cfl
    write(*, '(a, F7.3, a, F8.3)', advance='no') '  |', dg%trc_min, '< trC <', dg%trc_max
    if (method == 2) write(*, '(a, e12.4)', advance='no') '  | drag:', dg%drag
    write(*, *) ''
    if (dg%nan_flag /= 0) stop 'NaN value'
  end subroutine log_progress

  subroutine write_koide(s)
    integer, intent(in) :: s
    integer :: iu
    call check(fp_get_state(ctx, U, V, W, P, C, c_null_ptr), ctx, 'fp_get_state')
    open(newunit=iu, file=trim(dir)//'koide3.d', position='append')
    write(iu, *) s, C(1, NX/2, NY/2, NZ/2) + C(4, NX/2, NY/2, NZ/2) + C(6, NX/2, NY/2, NZ/2)
    close(iu)
  end subroutine write_koide

  subroutine write_time_stats(s)
    ! one record of Log/time_<key>.d (all_time_dif, fenep_mpi.f90:2342-2345),
    ! statistics reduced on the device; step of the pass (it restarts on the
    ! second pass of input_type = 2, as the reference's loop variable)
    integer, intent(in) :: s
    integer :: iu
    real(c_double) :: ts(13)
    call check(fp_time_stats(ctx, ts), ctx, 'fp_time_stats')
    call execute_command_line('mkdir -p '//trim(dir)//'Log')
    open(newunit=iu, file=trim(dir)//'Log/time_'//trim(key)//'.d', position='append')
    ! one line per record (list-directed output may wrap long records)
    write(iu, '(i10, 12es25.16e3)') s, s*dt, ts(3:13)
    close(iu)
  end subroutine write_time_stats

  subroutine add_energy()
    ! shell spectrum of the current state, accumulated between Gstep writes
    ! (fft_energy_dif, fenep_mpi.f90:2566-2665)
    if (.not. allocated(Energy)) then
      allocate(Energy(0:NX-1), Etmp(0:NX-1))
      Energy = 0.0d0
    endif
    call check(fp_energy_spectrum(ctx, Etmp, int(NX, c_int)), ctx, 'fp_energy_spectrum')
    Energy = Energy + Etmp
    n_energy = n_energy + 1
  end subroutine add_energy

  subroutine write_energy(s)
    ! Log/energy_same_<step>.d: the spectrum averaged over the accumulated
    ! records, rows k = 0..NX-1
    integer, intent(in) :: s
    integer :: iu, i
    character(8) :: str
    write(str, '(I8.8)') s
    call execute_command_line('mkdir -p '//trim(dir)//'Log')
    open(newunit=iu, file=trim(dir)//'Log/energy_'//trim(key)//'_'//str//'.d', status='replace')
    do i = 0, NX - 1
      write(iu, '(i8, es25.16e3)') i, Energy(i)/dble(n_energy)
    enddo
    close(iu)
    Energy = 0.0d0
    n_energy = 0
  end subroutine write_energy

  subroutine cell_values(i, j, k, qti, r)
    ! the 18 values of a map record at cell (i, j, k) (fenep.f90:996-1017,
    ! 1121-1146); column 7 holds Qti (get_data, get_data_xy, get_data_xz)
    ! or |u|^2/2 (get_data_binary)
    integer, intent(in) :: i, j, k
    logical, intent(in) :: qti
    real(8), intent(out) :: r(18)
    real(8) :: D(3, 3), Om(3), E(3)
    D(1, 1) = (U(i, j, k) - U(i-1, j, k))/dX
    D(1, 2) = (U(i, j+1, k) - U(i, j-1, k) + U(i-1, j+1, k) - U(i-1, j-1, k))/(4*dY)
    D(1, 3) = (U(i, j, k+1) - U(i, j, k-1) + U(i-1, j, k+1) - U(i-1, j, k-1))/(4*dZ)
    D(2, 1) = (V(i+1, j, k) - V(i-1, j, k) + V(i+1, j-1, k) - V(i-1, j-1, k))/(4*dX)
    D(2, 2) = (V(i, j, k) - V(i, j-1, k))/dY
    D(2, 3) = (V(i, j, k+1) - V(i, j, k-1) + V(i, j-1, k+1) - V(i, j-1, k-1))/(4*dZ)
    D(3, 1) = (W(i+1, j, k) - W(i-1, j, k) + W(i+1, j, k-1) - W(i-1, j, k-1))/(4*dX)
    D(3, 2) = (W(i, j+1, k) - W(i, j-1, k) + W(i, j+1, k-1) - W(i, j-1, k-1))/(4*dY)
    D(3, 3) = (W(i, j, k) - W(i, j, k-1))/dZ
    E(1) = (U(i-1, j, k) + U(i, j, k))/2
    E(2) = (V(i, j-1, k) + V(i, j, k))/2
    E(3) = (W(i, j, k-1) + W(i, j, k))/2
    Om(1) = D(3, 2) - D(2, 3)
    Om(2) = D(1, 3) - D(3, 1)
    Om(3) = D(2, 1) - D(1, 2)
    r(1) = (i - 0.5d0)*dX; r(2) = (j - 0.5d0)*dY; r(3) = (k - 0.5d0)*dZ
    r(4:6) = E
    if (qti) then
      r(7) = D(2, 2)*D(3, 3) - D(3, 2)*D(2, 3) + D(1, 1)*D(2, 2) - D(2, 1)*D(1, 2) + D(1, 1)*D(3, 3) - D(3, 1)*D(1, 3)
    else
      r(7) = sum(E**2)/2
    endif
    r(8:10) = Om
    r(11) = sum(Om**2)/2
    r(12) = C(1, i, j, k) + C(4, i, j, k) + C(6, i, j, k)
    r(13:18) = C(1:6, i, j, k)
  end subroutine cell_values

  subroutine plane_row(iu, i, j, k)
    ! one 18-column record of get_data_binary (fenep.f90:1121-1146)
    integer, intent(in) :: iu, i, j, k
    real(8) :: r(18)
    call cell_values(i, j, k, .false., r)
    write(iu) r
  end subroutine plane_row

  subroutine get_data_text(s)
    ! text maps of get_data_xy (k = NZ/2), get_data_xz (j = NY/2) and, for
    ! map_text = 2, get_data (every cell), format (18e12.4) (fenep.f90:982-1106)
    integer, intent(in) :: s
    character(8) :: str
    integer :: iu
    real(8) :: r(18)
    write(str, '(I8.8)') s
    open(newunit=iu, file=trim(dir)//'z_'//str//'.d', status='replace')
    k = NZ/2
    do j = 1, NY
      do i = 1, NX
        call cell_values(i, j, k, .true., r)
        write(iu, '(18e12.4)') r
      enddo
    enddo
    close(iu)
    open(newunit=iu, file=trim(dir)//'y_'//str//'.d', status='replace')
    j = NY/2
    do k = 1, NZ
      do i = 1, NX
        call cell_values(i, j, k, .true., r)
        write(iu, '(18e12.4)') r
      enddo
    enddo
    close(iu)
    if (map_text >= 2) then
      open(newunit=iu, file=trim(dir)//'all_'//str//'.d', status='replace')
      do k = 1, NZ
        do j = 1, NY
          do i = 1, NX
            call cell_values(i, j, k, .true., r)
            write(iu, '(18e12.4)') r
          enddo
        enddo
      enddo
      close(iu)
    endif
  end subroutine get_data_text

  subroutine get_data_binary(s)
    integer, intent(in) :: s
    character(8) :: str
    integer :: iu
    write(str, '(I8.8)') s
    call check(fp_get_state(ctx, U, V, W, P, C, c_null_ptr), ctx, 'fp_get_state')
    open(newunit=iu, file=trim(dir)//'Map/z_'//str//'.bin', form='unformatted', status='replace', access='stream')
    k = NZ/2
    do j = 1, NY
      do i = 1, NX
        call plane_row(iu, i, j, k)
      enddo
    enddo
    close(iu)
    open(newunit=iu, file=trim(dir)//'Map/y_'//str//'.bin', form='unformatted', status='replace', access='stream')
    j = NY/2
    do k = 1, NZ
      do i = 1, NX
        call plane_row(iu, i, j, k)
      enddo
    enddo
    close(iu)
    open(newunit=iu, file=trim(dir)//'Map/x_'//str//'.bin', form='unformatted', status='replace', access='stream')
    i = NX/2
    do k = 1, NZ
      do j = 1, NY
        call plane_row(iu, i, j, k)
      enddo
    enddo
    close(iu)
  end subroutine get_data_binary

  subroutine write_restart()
    integer :: iu
    call check(fp_get_state(ctx, U, V, W, P, C, c_loc(hist)), ctx, 'fp_get_state')
    open(newunit=iu, file=trim(dir)//'Restart.d', status='replace')
    do k = 1, NZ
      do j = 1, NY
        do i = 1, NX
          write(iu, '(16(es25.17e3, 1x))') U(i, j, k), V(i, j, k), W(i, j, k), P(i, j, k), C(:, i, j, k), &
               hist(i, j, k, 1), hist(i, j, k, 2), hist(i, j, k, 3), hist(i, j, k, 4), hist(i, j, k, 5), hist(i, j, k, 6)
        enddo
      enddo
    enddo
    close(iu)
  end subroutine write_restart

  subroutine read_restart()
    integer :: iu
    hist = 0.0d0
    open(newunit=iu, file=trim(dir)//'Restart.d', status='old', action='read')
    do k = 1, NZ
      do j = 1, NY
        do i = 1, NX
          read(iu, *) U(i, j, k), V(i, j, k), W(i, j, k), P(i, j, k), C(:, i, j, k), &
               hist(i, j, k, 1), hist(i, j, k, 2), hist(i, j, k, 3), hist(i, j, k, 4), hist(i, j, k, 5), hist(i, j, k, 6)
        enddo
      enddo
    enddo
    close(iu)
  end subroutine read_restart

end program fenep_gpu
