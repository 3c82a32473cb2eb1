!! CUBE-style Fortran driver over libcubehip.so.
!!
!! Same program shape as CUBE's main (main/main.f90:19-100): initialise,
!! read the checkpoint at z_checkpoint(1), then loop timestep -> update_x ->
!! [ext_pp_force] -> particle_mesh, writing a checkpoint (after the half
!! drift with dt_old=0) at each redshift of z_checkpoint.txt and running the
!! halo finder at those of z_halofind.txt (HALOFIND).  The step
!! routines are the library's entry points; the time stepping (timestep.f90)
!! stays on the host in real(4) as in the reference.
!!
!! Run-time parameters come from the namelist file given as the first
!! argument (default cube.nml), group /cube/, with CUBE's parameter names.
program cube_main
  use iso_c_binding
  use cube_hip
  implicit none

  ! parameters.f90 names, set at run time
  integer(8) :: nc = 32, nnt = 2, ncb = 6, izipx = 2, izipv = 2, istep_max = 2500
  logical :: Extended_pp_force = .false.
  integer :: pp_range = 1, device = 0
  real :: rsoft = 0.3, omega_m = 0.31, omega_l = 0.69, wde = -1, box = 100, h0 = 0.68
  real, parameter :: ra_max = 0.2, dt_max = 1
  character(256) :: opath = 'output/', kernel_dir = 'kernels/', zfile = 'z_checkpoint.txt'
  ! HALOFIND: hfile names z_halofind.txt ('' leaves the halo finder off);
  ! halofind.f90:36-40 constants, nq the pid lattice size (0: nf)
  character(256) :: hfile = ''
  real :: den_peak_cutoff = 100
  integer :: min_halo_particles = 100
  integer(8) :: nq = 0
  namelist /cube/ nc, nnt, ncb, izipx, izipv, istep_max, Extended_pp_force, pp_range, rsoft, &
                  omega_m, omega_l, wde, box, h0, opath, kernel_dir, zfile, device, &
                  hfile, den_peak_cutoff, min_halo_particles, nq

  type(cube_params) :: p
  type(cube_sim_header) :: sim
  type(cube_drift_out) :: dro
  type(cube_pp_out) :: ppo
  type(cube_pm_out) :: pmo
  type(c_ptr) :: h
  integer(c_int) :: rc
  real(c_float), allocatable :: kf(:), kc(:)
  integer(c_int8_t), allocatable, target :: xpb(:), vpb(:)
  integer(c_int32_t), allocatable :: pid(:), rhoc(:)
  real(c_float), allocatable :: vfield(:)
  real :: z_checkpoint(1000), z_halofind(1000)
  integer :: n_checkpoint, cur_checkpoint, n_halofind, cur_halofind, i, u
  integer(8) :: istep, np_max
  real :: a, a_mid, da, dt, dt_old, dt_neu, t, tau
  logical :: checkpoint_step, final_step, halofind_step, halofind_on
  character(256) :: nml

  nml = 'cube.nml'
  if (command_argument_count() >= 1) call get_command_argument(1, nml)
  open(newunit=u, file=trim(nml), status='old', action='read')
  read(u, nml=cube)
  close(u)

  ! initialize.f90:55-61
  open(newunit=u, file=trim(zfile), status='old', action='read')
  n_checkpoint = 0
  do i = 1, 999
    read(u, end=71, fmt='(f8.4)') z_checkpoint(i)
    n_checkpoint = i
  enddo
71 close(u)
  if (n_checkpoint == 0) error stop 'z_checkpoint.txt empty'
  ! initialize.f90:62-71; entries past the list stay 0 as in the reference
  z_halofind = 0
  n_halofind = 0
  halofind_on = len_trim(hfile) > 0
  if (halofind_on) then
    open(newunit=u, file=trim(hfile), status='old', action='read')
    do i = 1, 999
      read(u, end=81, fmt='(f8.4)') z_halofind(i)
      n_halofind = i
    enddo
81  close(u)
    if (n_halofind == 0) error stop 'z_halofind.txt empty'
  endif

  ! kernel tables (kernel_f.f90:14-23, kernel_c.f90:31-39)
  allocate(kf(4096 * 3), kc(64 * 3))
  call read_table(trim(kernel_dir)//'wfxyzf.3.ascii', 4096, kf)
  call read_table(trim(kernel_dir)//'wfxyzc.2.ascii', 64, kc)

  ! particle_initialization.f90:34-82
  open(newunit=u, file=output_name(z_checkpoint(1), 'info'), status='old', access='stream')
  read(u) sim
  close(u)
  if (sim%izipx /= izipx .or. sim%izipv /= izipv) error stop 'zip format incompatable'
  np_max = max(sim%nplocal, (2 * nc)**3)
  allocate(xpb(3 * np_max * izipx), vpb(3 * np_max * izipv), pid(np_max), rhoc(nc**3), vfield(3 * nc**3))
  call read_bytes(output_name(z_checkpoint(1), 'xp'), xpb, 3 * sim%nplocal * izipx)
  call read_bytes(output_name(z_checkpoint(1), 'vp'), vpb, 3 * sim%nplocal * izipv)
  open(newunit=u, file=output_name(z_checkpoint(1), 'np'), status='old', access='stream')
  read(u) rhoc
  close(u)
  open(newunit=u, file=output_name(z_checkpoint(1), 'vc'), status='old', access='stream')
  read(u) vfield
  close(u)
  open(newunit=u, file=output_name(z_checkpoint(1), 'id'), status='old', access='stream')
  read(u) pid(1:sim%nplocal)
  close(u)

  p%nc = nc; p%nnt = nnt; p%ncb = ncb; p%izipx = izipx; p%izipv = izipv; p%np_max = np_max
  p%pp_on = merge(1, 0, Extended_pp_force); p%pp_range = pp_range; p%rsoft = rsoft
  p%omega_m = omega_m; p%omega_l = omega_l; p%wde = wde; p%precision = 4; p%device = device
  rc = cube_init(p, kf, kc, h)
  call cube_ok(h, rc, 'cube_init')
  rc = cube_upload_state(h, sim, c_loc(xpb), c_loc(vpb), pid, rhoc, vfield)
  call cube_ok(h, rc, 'cube_upload_state')

  ! initialize.f90:40-53, particle_initialization.f90:37-38, main.f90:42
  a = 1. / (1 + z_checkpoint(1))
  t = 0; dt = 0; dt_old = 0; da = 0; dt_neu = dt_max
  tau = -3 / sqrt(1. / (1 + z_checkpoint(1)))
  cur_checkpoint = 2
  cur_halofind = cur_checkpoint + 1   ! main.f90:43
  checkpoint_step = .false.; final_step = .false.; halofind_step = .false.
  print*, 'cube_main: nplocal =', sim%nplocal, ' z_i =', z_checkpoint(1)

  do istep = 1, istep_max
    call timestep
    rc = cube_drift(h, dt_old, dt, dro)
    call cube_ok(h, rc, 'cube_drift')
    if (Extended_pp_force) then
      rc = cube_pp(h, a_mid, dt, ppo)
      call cube_ok(h, rc, 'cube_pp')
      sim%dt_pp = ppo%dt_pp
    endif
    rc = cube_pm(h, a_mid, dt, pmo)
    call cube_ok(h, rc, 'cube_pm')
    sim%dt_fine = pmo%dt_fine; sim%dt_coarse = pmo%dt_coarse; sim%dt_vmax = pmo%dt_vmax
    print '(a,i5,a,f9.5,a,es12.5,a,es12.5)', ' step', istep, '  a=', a, '  dt=', dt, '  sigma_vi=', dro%sigma_vi_new
    if (checkpoint_step .or. halofind_step) then   ! main.f90:72-95
      dt_old = 0
      rc = cube_drift(h, dt_old, dt, dro)
      call cube_ok(h, rc, 'cube_drift')
      if (checkpoint_step) then
        call checkpoint(z_checkpoint(cur_checkpoint))
        cur_checkpoint = cur_checkpoint + 1
      endif
      if (halofind_step) then
        call halofind(z_halofind(cur_halofind))
        cur_halofind = cur_halofind + 1
      endif
      if (final_step) exit
      dt = 0
    endif
  enddo
  rc = cube_finalize(h)
  print*, 'cube_main: done after', istep, 'steps'

contains

  !! timestep.f90:8-95 on one image
  subroutine timestep
    real :: dt_e, ra, da1, da2, a_next, z_next
    integer :: ntry, irc
    dt_old = dt
    sim%timestep = sim%timestep + 1
    dt_e = dt_max
    ntry = 0
    do
      ntry = ntry + 1
      irc = cube_expansion(a, dt_e, omega_m, omega_l, wde, da1, da2)
      da = da1 + da2
      ra = da / (a + da)
      if (ra > ra_max) then
        dt_e = dt_e * (ra_max / ra)
      else
        exit
      endif
      if (ntry > 10) exit
    enddo
    dt = min(dt_e, sim%dt_fine, sim%dt_coarse, sim%dt_pp, sim%dt_vmax, 1000., dt_neu)
    irc = cube_expansion(a, dt, omega_m, omega_l, wde, da1, da2)
    da = da1 + da2
    checkpoint_step = .false.
    halofind_step = .false.
    z_next = z_checkpoint(cur_checkpoint)
    if (halofind_on) z_next = max(z_checkpoint(cur_checkpoint), z_halofind(cur_halofind))
    a_next = 1.0 / (1 + z_next)
    if (da >= a_next - a) then
      if (z_next == z_checkpoint(cur_checkpoint)) then
        checkpoint_step = .true.
        if (cur_checkpoint == n_checkpoint) final_step = .true.
      endif
      if (halofind_on .and. z_next == z_halofind(cur_halofind)) halofind_step = .true.
      do while (abs((a + da) / a_next - 1) >= 1e-6)
        dt = dt * (a_next - a) / da
        irc = cube_expansion(a, dt, omega_m, omega_l, wde, da1, da2)
        da = da1 + da2
      enddo
    endif
    a_mid = a + (da / 2)
    tau = tau + dt
    t = t + dt
    a = a + da
    dt_neu = 1000.
  end subroutine

  !! checkpoint.f90:24-47
  subroutine checkpoint(z)
    real, intent(in) :: z
    integer :: v, irc
    integer(c_int64_t) :: nstep
    nstep = sim%timestep   ! counted here, in timestep (timestep.f90:10)
    irc = cube_download_state(h, sim, c_loc(xpb), c_loc(vpb), pid, rhoc, vfield)
    call cube_ok(h, irc, 'cube_download_state')
    sim%timestep = nstep
    sim%a = a; sim%t = t; sim%tau = tau
    sim%cur_checkpoint = cur_checkpoint
    sim%vsim2phys = (1.5 / a) * box * 100. * sqrt(omega_m) / (nc * 4)
    call execute_command_line('mkdir -p '//trim(opath)//'image1')
    open(newunit=v, file=output_name(z, 'info'), status='replace', access='stream'); write(v) sim; close(v)
    open(newunit=v, file=output_name(z, 'xp'), status='replace', access='stream')
    write(v) xpb(1:3 * sim%nplocal * izipx); close(v)
    open(newunit=v, file=output_name(z, 'vp'), status='replace', access='stream')
    write(v) vpb(1:3 * sim%nplocal * izipv); close(v)
    open(newunit=v, file=output_name(z, 'np'), status='replace', access='stream'); write(v) rhoc; close(v)
    open(newunit=v, file=output_name(z, 'vc'), status='replace', access='stream'); write(v) vfield; close(v)
    open(newunit=v, file=output_name(z, 'id'), status='replace', access='stream')
    write(v) pid(1:sim%nplocal); close(v)
    print*, 'checkpoint written at z =', z, ' nplocal =', sim%nplocal
  end subroutine

  !! halofind.f90 through cube_halofind; writes the halo, halo_pid and hid
  !! files of halofind.f90:117-125,568-578
  subroutine halofind(z)
    real, intent(in) :: z
    type(cube_halo_params) :: hp
    type(cube_halo_catalog) :: cat
    real(c_float), allocatable, target :: halos(:, :)
    integer(c_int32_t), allocatable, target :: words(:)
    integer(c_int16_t), allocatable, target :: hid(:)
    integer(8) :: maxh, maxw, nf
    integer :: v, irc
    nf = 4 * nc
    maxh = 1048576
    maxw = 2 * maxh + sim%nplocal + 16
    allocate(halos(42, maxh), words(maxw), hid(nf**3))
    hp%den_odc = 200; hp%den_peak_cutoff = den_peak_cutoff; hp%min_halo_particles = min_halo_particles
    hp%nc_halo_max = 128; hp%newbox_max = 300; hp%search_ratio = 2; hp%refine_ratio = 3; hp%pad = 0
    hp%max_halo_np = 5 * 129**3 / 3
    hp%nq = merge(nq, nf, nq > 0)
    hp%h0 = h0; hp%box = box; hp%los_centre = [30., 370., 370.]
    hp%a = a
    hp%vsim2phys = (1.5 / a) * box * 100. * sqrt(omega_m) / nf   ! checkpoint.f90:30
    hp%pad2 = 0
    cat%max_halos = maxh; cat%halos = c_loc(halos)
    cat%max_pid_words = maxw; cat%pid_words = c_loc(words)
    cat%nhid = nf**3; cat%hid = c_loc(hid)
    irc = cube_halofind(h, hp, cat)
    call cube_ok(h, irc, 'cube_halofind')
    call execute_command_line('mkdir -p '//trim(opath)//'image1')
    open(newunit=v, file=output_name(z, 'halo'), status='replace', access='stream')
    write(v) int(cat%nhalo, c_int32_t), int(cat%nhalo, c_int32_t), hp%den_odc
    write(v) halos(:, 1:cat%nhalo)
    close(v)
    open(newunit=v, file=output_name(z, 'halo_pid'), status='replace', access='stream')
    write(v) words(1:cat%npid_words)
    close(v)
    open(newunit=v, file=output_name(z, 'hid'), status='replace', access='stream')
    write(v) hid
    close(v)
    print*, 'halofind at z =', z, ': nhalo =', cat%nhalo
    deallocate(halos, words, hid)
  end subroutine

  !! basic_functions.fh:52-87
  function output_name(z, what) result(fn)
    real, intent(in) :: z
    character(*), intent(in) :: what
    character(:), allocatable :: fn
    character(20) :: zs
    write(zs, '(f7.3)') z
    fn = trim(opath)//'image1/'//trim(adjustl(zs))//'_'//what//'_1.bin'
  end function

  subroutine read_table(fn, nrow, tab)
    character(*), intent(in) :: fn
    integer, intent(in) :: nrow
    real(c_float), intent(out) :: tab(3, nrow)
    integer :: v, r, idx(3)
    open(newunit=v, file=fn, status='old', action='read')
    do r = 1, nrow
      read(v, *) idx, tab(:, r)
    enddo
    close(v)
  end subroutine

  subroutine read_bytes(fn, buf, n)
    character(*), intent(in) :: fn
    integer(c_int8_t), intent(out) :: buf(*)
    integer(8), intent(in) :: n
    integer :: v
    open(newunit=v, file=fn, status='old', access='stream')
    read(v) buf(1:n)
    close(v)
  end subroutine

end program cube_main
