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
