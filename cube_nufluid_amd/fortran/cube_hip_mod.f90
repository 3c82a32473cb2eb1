!! iso_c_binding interface of libcubehip.so (include/cube_hip.h) for a
!! CUBE-style Fortran driver.  Every call returns an integer status; cube_ok
!! turns a nonzero status into 'error stop' with the library's message, as the
!! reference's fail-stop checks do (e.g. update_particle.f90:90-96).
module cube_hip
  use iso_c_binding
  implicit none

  type, bind(C) :: cube_params
    integer(c_int64_t) :: nc, nnt, ncb, izipx, izipv, np_max
    integer(c_int32_t) :: pp_on, pp_range
    real(c_float) :: rsoft, omega_m, omega_l, wde
    integer(c_int32_t) :: precision, device
  end type

  !! parameters.f90:129-153, 224 bytes
  type, bind(C) :: cube_sim_header
    integer(c_int64_t) :: nplocal, npglobal, nplocal_nu, npglobal_nu
    integer(c_int64_t) :: izipx, izipv, izipx_nu, izipv_nu
    integer(c_int64_t) :: image
    integer(c_int64_t) :: nn, nnt, nt, ncell, ncb
    integer(c_int64_t) :: timestep, cur_checkpoint, cur_halofind
    real(c_float) :: a, t, tau
    real(c_float) :: dt_pp, dt_fine, dt_coarse, dt_vmax, dt_vmax_nu
    real(c_float) :: mass_p_cdm, mass_p_nu
    real(c_float) :: box
    real(c_float) :: h0, omega_m, omega_l, s8, vsim2phys, sigma_vres, sigma_vi, sigma_vi_nu
    real(c_float) :: z_i, z_i_nu
    real(c_float) :: vz_max
  end type

  type, bind(C) :: cube_drift_out
    real(c_float) :: sigma_vi_new
    integer(c_int32_t) :: nlayer
    real(c_double) :: std_vsim, std_vsim_c, std_vsim_res
    integer(c_int64_t) :: npcheck
    integer(c_int32_t) :: max_disp(3)
    integer(c_int32_t) :: frame_conflicts
  end type

  type, bind(C) :: cube_pp_out
    real(c_float) :: dt_pp, f2_max_pp
    integer(c_int64_t) :: npairs, itest1, dense_blocks
  end type

  type, bind(C) :: cube_pm_out
    real(c_float) :: dt_fine, dt_coarse, dt_vmax, vz_max
    real(c_float) :: vmax(3)
    real(c_float) :: f2_max_fine, f2_max_coarse
    real(c_double) :: mass_f, mass_c
  end type

  !! halofind.f90:27-40 constants and the header values it reads (80 bytes)
  type, bind(C) :: cube_halo_params
    real(c_float) :: den_odc, den_peak_cutoff
    integer(c_int32_t) :: min_halo_particles, nc_halo_max, newbox_max, search_ratio, refine_ratio, pad
    integer(c_int64_t) :: max_halo_np, nq
    real(c_float) :: h0, box, los_centre(3), a, vsim2phys, pad2
  end type

  !! host buffers of the halo, halo_pid and hid outputs (80 bytes)
  type, bind(C) :: cube_halo_catalog
    integer(c_int64_t) :: max_halos
    type(c_ptr) :: halos
    integer(c_int64_t) :: max_pid_words
    type(c_ptr) :: pid_words
    integer(c_int64_t) :: nhid
    type(c_ptr) :: hid
    integer(c_int64_t) :: nhalo, npid_words, npeaks, nprocessed
  end type

  !! neutrino-fluid set-up, one neutrino mass (32 bytes)
  type, bind(C) :: cube_nu_params
    real(c_float) :: m_ev, f_neu, z_i_nu, gamma, box, h0, omega_m
    integer(c_int32_t) :: neutrino_ic
  end type

  !! seeded Zel'dovich initial conditions; lnk and lnp point to nk doubles
  !! (56 bytes)
  type, bind(C) :: cube_ic_params
    integer(c_int64_t) :: seed
    integer(c_int32_t) :: np_nc, nk
    real(c_float) :: a, box
    real(c_double) :: vfac
    type(c_ptr) :: lnk, lnp
    real(c_float) :: f_cdm
    integer(c_int32_t) :: reserved
  end type

  !! time-stepping state of the native host loop (320 bytes)
  integer, parameter :: cube_max_checkpoints = 64
  type, bind(C) :: cube_clock
    real(c_float) :: a, t, tau, dt, dt_old, da, a_mid, dt_neu
    integer(c_int32_t) :: cur_checkpoint, n_checkpoint, checkpoint_step, final_step
    integer(c_int32_t) :: checkpoint_at, pad
    integer(c_int64_t) :: istep
    real(c_float) :: z_checkpoint(cube_max_checkpoints)
  end type

  !! ensemble engine parameters (24 bytes) and one finished run (192 bytes)
  type, bind(C) :: cube_ens_params
    integer(c_int32_t) :: device, max_live
    integer(c_int64_t) :: mem_budget
    integer(c_int32_t) :: step_mode, reserved
  end type
  type, bind(C) :: cube_ens_result
    integer(c_int64_t) :: id
    integer(c_int32_t) :: status, steps
    real(c_double) :: t_submit, t_admit, t_done
    integer(c_int64_t) :: nparticles
    real(c_float) :: a_final
    integer(c_int32_t) :: pad
    integer(c_int64_t) :: state_hash
    character(kind=c_char) :: message(128)
  end type

  interface
    integer(c_int) function cube_init(p, kf, kc, h) bind(C, name="cube_init")
      import :: c_int, c_float, c_ptr, cube_params
      type(cube_params), intent(in) :: p
      real(c_float), intent(in) :: kf(*), kc(*)
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function cube_upload_state(h, sim, xp, vp, pid, rhoc, vfield) bind(C, name="cube_upload_state")
      import :: c_int, c_ptr, c_int32_t, c_float, cube_sim_header
      type(c_ptr), value :: h
      type(cube_sim_header), intent(in) :: sim
      type(c_ptr), value :: xp, vp
      integer(c_int32_t), intent(in) :: pid(*), rhoc(*)
      real(c_float), intent(in) :: vfield(*)
    end function
    integer(c_int) function cube_download_state(h, sim, xp, vp, pid, rhoc, vfield) bind(C, name="cube_download_state")
      import :: c_int, c_ptr, c_int32_t, c_float, cube_sim_header
      type(c_ptr), value :: h
      type(cube_sim_header), intent(out) :: sim
      type(c_ptr), value :: xp, vp
      integer(c_int32_t), intent(out) :: pid(*), rhoc(*)
      real(c_float), intent(out) :: vfield(*)
    end function
    integer(c_int) function cube_drift(h, dt_old, dt, out) bind(C, name="cube_drift")
      import :: c_int, c_ptr, c_float, cube_drift_out
      type(c_ptr), value :: h
      real(c_float), value :: dt_old, dt
      type(cube_drift_out), intent(out) :: out
    end function
    integer(c_int) function cube_pp(h, a_mid, dt, out) bind(C, name="cube_pp")
      import :: c_int, c_ptr, c_float, cube_pp_out
      type(c_ptr), value :: h
      real(c_float), value :: a_mid, dt
      type(cube_pp_out), intent(out) :: out
    end function
    integer(c_int) function cube_pm(h, a_mid, dt, out) bind(C, name="cube_pm")
      import :: c_int, c_ptr, c_float, cube_pm_out
      type(c_ptr), value :: h
      real(c_float), value :: a_mid, dt
      type(cube_pm_out), intent(out) :: out
    end function
    integer(c_int) function cube_expansion(a, dt, om, ol, wde, da1, da2) bind(C, name="cube_expansion")
      import :: c_int, c_float
      real(c_float), value :: a, dt, om, ol, wde
      real(c_float), intent(out) :: da1, da2
    end function
    integer(c_int) function cube_last_error(h, buf, n) bind(C, name="cube_last_error")
      import :: c_int, c_ptr, c_char, c_size_t
      type(c_ptr), value :: h
      character(kind=c_char), intent(out) :: buf(*)
      integer(c_size_t), value :: n
    end function
    integer(c_int) function cube_halofind(h, hp, cat) bind(C, name="cube_halofind")
      import :: c_int, c_ptr, cube_halo_params, cube_halo_catalog
      type(c_ptr), value :: h
      type(cube_halo_params), intent(in) :: hp
      type(cube_halo_catalog), intent(inout) :: cat
    end function
    integer(c_int) function cube_finalize(h) bind(C, name="cube_finalize")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function cube_ic_zeldovich(h, ip, sim_out) bind(C, name="cube_ic_zeldovich")
      import :: c_int, c_ptr, cube_ic_params, cube_sim_header
      type(c_ptr), value :: h
      type(cube_ic_params), intent(in) :: ip
      type(cube_sim_header), intent(out) :: sim_out
    end function
    integer(c_int) function cube_nu_init(h, p) bind(C, name="cube_nu_init")
      import :: c_int, c_ptr, cube_nu_params
      type(c_ptr), value :: h
      type(cube_nu_params), intent(in) :: p
    end function
    integer(c_int) function cube_nu_ic(h, ip) bind(C, name="cube_nu_ic")
      import :: c_int, c_ptr, cube_ic_params
      type(c_ptr), value :: h
      type(cube_ic_params), intent(in) :: ip
    end function
    integer(c_int) function cube_nu_info(h, n, nc, dx, cs2, weight) bind(C, name="cube_nu_info")
      import :: c_int, c_ptr, c_int32_t, c_float
      type(c_ptr), value :: h
      integer(c_int), value :: n
      integer(c_int32_t), intent(out) :: nc
      real(c_float), intent(out) :: dx, cs2, weight
    end function
    integer(c_int) function cube_nu_set_fluid(h, n, fld) bind(C, name="cube_nu_set_fluid")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: h
      integer(c_int), value :: n
      real(c_float), intent(in) :: fld(*)
    end function
    integer(c_int) function cube_nu_get_fluid(h, n, fld) bind(C, name="cube_nu_get_fluid")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: h
      integer(c_int), value :: n
      real(c_float), intent(out) :: fld(*)
    end function
    integer(c_int) function cube_fluid_evolve(h, dt, dir, a, dt_neu) bind(C, name="cube_fluid_evolve")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: h
      real(c_float), value :: dt
      integer(c_int), value :: dir
      real(c_float), value :: a
      real(c_float), intent(inout) :: dt_neu
    end function
    integer(c_int) function cube_clock_init(c, z_checkpoint, n) bind(C, name="cube_clock_init")
      import :: c_int, c_float, cube_clock
      type(cube_clock), intent(out) :: c
      real(c_float), intent(in) :: z_checkpoint(*)
      integer(c_int), value :: n
    end function
    integer(c_int) function cube_step(h, c) bind(C, name="cube_step")
      import :: c_int, c_ptr, cube_clock
      type(c_ptr), value :: h
      type(cube_clock), intent(inout) :: c
    end function
    integer(c_int) function cube_set_step_mode(h, mode) bind(C, name="cube_set_step_mode")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: mode
    end function
    integer(c_int) function cube_step_stats(h, graph_instantiations, graph_updates) bind(C, name="cube_step_stats")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: h
      integer(c_int64_t), intent(out) :: graph_instantiations, graph_updates
    end function
    integer(c_int64_t) function cube_footprint(p, ic_np_nc) bind(C, name="cube_footprint")
      import :: c_int64_t, c_int32_t, cube_params
      type(cube_params), intent(in) :: p
      integer(c_int32_t), value :: ic_np_nc
    end function
    integer(c_int) function cube_power_spectrum(h, nbin, modes, ksum, delta2sum) bind(C, name="cube_power_spectrum")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: nbin
      real(c_double), intent(out) :: modes(*), ksum(*), delta2sum(*)
    end function
    integer(c_int) function cube_power_spectrum_noalloc(h, nbin, modes, ksum, delta2sum) &
        bind(C, name="cube_power_spectrum_noalloc")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: nbin
      real(c_double), intent(out) :: modes(*), ksum(*), delta2sum(*)
    end function
    integer(c_int) function cube_probe_density(h, rho) bind(C, name="cube_probe_density")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: h
      real(c_float), intent(out) :: rho(*)
    end function
    integer(c_int) function cube_probe_pm_forces(h, f_fine, f_coarse) bind(C, name="cube_probe_pm_forces")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: f_fine(*), f_coarse(*)
    end function
    integer(c_int) function cube_probe_pp_forces(h, f_pp) bind(C, name="cube_probe_pp_forces")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: f_pp(*)
    end function
    integer(c_int) function cube_probe_vencode(h, x, n, codes) bind(C, name="cube_probe_vencode")
      import :: c_int, c_ptr, c_double, c_int64_t, c_int32_t
      type(c_ptr), value :: h
      real(c_double), intent(in) :: x(*)
      integer(c_int64_t), value :: n
      integer(c_int32_t), intent(out) :: codes(*)
    end function
    integer(c_int) function cube_profile_enable(h, on) bind(C, name="cube_profile_enable")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: on
    end function
    integer(c_int) function cube_profile_read(h, ms, count, nslot) bind(C, name="cube_profile_read")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: h
      real(c_double), intent(out) :: ms(*)
      integer(c_int64_t), intent(out) :: count(*)
      integer(c_int), value :: nslot
    end function
    integer(c_int) function cube_profile_slot_name(slot, buf, n) bind(C, name="cube_profile_slot_name")
      import :: c_int, c_char, c_size_t
      integer(c_int), value :: slot
      character(kind=c_char), intent(out) :: buf(*)
      integer(c_size_t), value :: n
    end function
    integer(c_int) function cube_hbm_copy_probe(device, bytes, iters, gbps) bind(C, name="cube_hbm_copy_probe")
      import :: c_int, c_int64_t, c_double
      integer(c_int), value :: device
      integer(c_int64_t), value :: bytes
      integer(c_int), value :: iters
      real(c_double), intent(out) :: gbps(*)
    end function
    integer(c_int) function cube_ens_create(p, e) bind(C, name="cube_ens_create")
      import :: c_int, c_ptr, cube_ens_params
      type(cube_ens_params), intent(in) :: p
      type(c_ptr), intent(out) :: e
    end function
    integer(c_int) function cube_ens_submit(e, p, kf, kc, ic, z_checkpoint, n_z, id) bind(C, name="cube_ens_submit")
      import :: c_int, c_ptr, c_float, c_int64_t, cube_params, cube_ic_params
      type(c_ptr), value :: e
      type(cube_params), intent(in) :: p
      real(c_float), intent(in) :: kf(*), kc(*)
      type(cube_ic_params), intent(in) :: ic
      real(c_float), intent(in) :: z_checkpoint(*)
      integer(c_int), value :: n_z
      integer(c_int64_t), intent(out) :: id
    end function
    integer(c_int) function cube_ens_poll(e, out, max_results, n_out) bind(C, name="cube_ens_poll")
      import :: c_int, c_ptr, cube_ens_result
      type(c_ptr), value :: e
      type(cube_ens_result), intent(out) :: out(*)
      integer(c_int), value :: max_results
      integer(c_int), intent(out) :: n_out
    end function
    integer(c_int) function cube_ens_pending(e, queued, live) bind(C, name="cube_ens_pending")
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), intent(out) :: queued, live
    end function
    real(c_double) function cube_ens_clock(e) bind(C, name="cube_ens_clock")
      import :: c_double, c_ptr
      type(c_ptr), value :: e
    end function
    integer(c_int) function cube_ens_destroy(e) bind(C, name="cube_ens_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: e
    end function
  end interface

contains

  subroutine cube_ok(h, rc, what)
    type(c_ptr), intent(in) :: h
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char) :: buf(512)
    character(512) :: msg
    integer :: i, irc
    if (rc == 0) return
    irc = cube_last_error(h, buf, int(512, c_size_t))
    msg = ''
    do i = 1, 512
      if (buf(i) == c_null_char) exit
      msg(i:i) = buf(i)
    enddo
    print*, 'error in ', what, ' (status', rc, '): ', trim(msg)
    error stop 1
  end subroutine

end module cube_hip
