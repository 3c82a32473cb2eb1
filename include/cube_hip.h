/*
 * cube_hip.h -- C ABI of libcubehip.so, the MI355X (gfx950) implementation of
 * CUBE's per-step gravity and update path.
 *
 * CUBE's driver (CUBE_v2.0/work/main/main.f90:47-98) calls argument-less
 * Fortran subroutines that share module-global state.  This library keeps
 * that state resident in HBM behind an opaque handle and exposes one entry
 * point per subroutine of the path.  Every entry point returns 0 on success
 * or a nonzero cube_status; the message is available from cube_last_error().
 * The Fortran binding (cube_nufluid_amd/fortran/cube_hip_mod.f90) turns a
 * nonzero status into 'error stop', matching the reference's fail-stop
 * 'stop' statements.
 *
 * Conventions: one host thread per handle; calls are ordered on the handle's
 * HIP stream; entry points that return scalars synchronise that stream.
 * Host arrays use CUBE's checkpoint layout (main/checkpoint.f90:36-47):
 * particles in tile order (tile x fastest), then coarse cell (x fastest),
 * then in-cell order; rhoc/vfield over physical cells in the same order.
 */
#ifndef CUBE_HIP_H
#define CUBE_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cube_handle cube_handle;

enum cube_status {
  CUBE_OK = 0,
  CUBE_ERR_ARG = 1,       /* invalid parameter or array size */
  CUBE_ERR_CAPACITY = 2,  /* particle capacity / displacement bound exceeded
                             (update_particle.f90:90-96, buffer_grid.f90:131-137) */
  CUBE_ERR_FORMAT = 3,    /* header/zip mismatch (particle_initialization.f90:41-45) */
  CUBE_ERR_CHECK = 4,     /* conservation check failed (ext_pp_force.f90:154-158) */
  CUBE_ERR_HIP = 5        /* HIP runtime error */
};

/* Compile-time constants of main/parameters.f90 (lines 8-116) that shape the
   path, as run-time values.  nn=1 (one image). */
typedef struct cube_params {
  int64_t nc;        /* coarse cells per dim (parameters.f90:29)             */
  int64_t nnt;       /* tiles per dim (:28); nt = nc/nnt                     */
  int64_t ncb;       /* coarse buffer cells (:47), 5 <= ncb <= nt            */
  int64_t izipx;     /* bytes per position component, 1 or 2 (:8)           */
  int64_t izipv;     /* bytes per velocity component, 1 or 2 (:9)           */
  int64_t np_max;    /* particle capacity of the resident arrays            */
  int32_t pp_on;     /* Extended_pp_force (:60)                             */
  int32_t pp_range;  /* (:62)                                               */
  float rsoft;       /* PP softening (:61)                                   */
  float omega_m;     /* (:91)                                               */
  float omega_l;     /* (:99)                                               */
  float wde;         /* (:100)                                              */
  int32_t precision; /* 4: FP32 meshes/forces (production); 8: FP64 oracle   */
  int32_t device;    /* HIP device ordinal                                   */
} cube_params;

/* 224-byte checkpoint header, main/parameters.f90:129-153 (17 x int64 then
   22 x float32, stream-written by checkpoint.f90:36). */
typedef struct cube_sim_header {
  int64_t nplocal, npglobal, nplocal_nu, npglobal_nu;
  int64_t izipx, izipv, izipx_nu, izipv_nu;
  int64_t image;
  int64_t nn, nnt, nt, ncell, ncb;
  int64_t timestep;
  int64_t cur_checkpoint;
  int64_t cur_halofind;
  float a, t, tau;
  float dt_pp, dt_fine, dt_coarse, dt_vmax, dt_vmax_nu;
  float mass_p_cdm, mass_p_nu;
  float box;
  float h0, omega_m, omega_l, s8, vsim2phys, sigma_vres, sigma_vi, sigma_vi_nu;
  float z_i, z_i_nu;
  float vz_max;
} cube_sim_header;

/* Results of one drift (update_particle.f90:221-252). */
typedef struct cube_drift_out {
  float sigma_vi_new;    /* std_vsim_res/sqrt(3.) (:226)                      */
  int32_t nlayer;        /* 2*ceiling(dt_mid*vz_max/ncell)+1 (:39)           */
  double std_vsim, std_vsim_c, std_vsim_res; /* (:221-223), simulation units */
  int64_t npcheck;       /* particles after re-binning (:247-252)            */
  int32_t max_disp[3];   /* largest cell displacement per dimension          */
  int32_t frame_conflicts; /* tile-frame rounding disagreements (expected 0) */
} cube_drift_out;

/* Results of one PP pass (ext_pp_force.f90:139-151). */
typedef struct cube_pp_out {
  float dt_pp;           /* 1/(sqrt(f2_max_pp)*a_mid*GG), running max (:139) */
  float f2_max_pp;
  int64_t npairs;        /* pair evaluations incl. self (:98)                */
  int64_t itest1;        /* particles updated (:87)                          */
  int64_t dense_blocks;  /* blocks too crowded for the LDS path (diagnostic)  */
} cube_pp_out;

/* Results of one particle_mesh call (pm.f90:427-446). */
typedef struct cube_pm_out {
  float dt_fine, dt_coarse, dt_vmax, vz_max;
  float vmax[3];
  float f2_max_fine, f2_max_coarse;
  double mass_f, mass_c;  /* testrhof, testrhoc (pm.f90:446)                */
} cube_pm_out;

/* Replaces initialize.f90:21-38 (create_cubefft_plan, create_penfft_plan,
   kernel_f, kernel_c).  kern_f_table: kernels/wfxyzf.3.ascii as 4096 rows x
   3 floats in file order (kernel_f.f90:14-23); kern_c_table:
   kernels/wfxyzc.2.ascii as 64 rows x 3 floats (kernel_c.f90:31-39). */
int cube_init(const cube_params* p, const float* kern_f_table, const float* kern_c_table,
              cube_handle** out);

/* Replaces particle_initialization.f90:34-82 (the reads) plus the halo
   set-up of buffer_grid/buffer_x/buffer_v (main.f90:20-23).  xp/vp hold
   3*nplocal components of izipx/izipv bytes; rhoc nc^3 int32; vfield
   3*nc^3 float32. */
int cube_upload_state(cube_handle* h, const cube_sim_header* sim, const void* xp, const void* vp,
                      const int32_t* pid, const int32_t* rhoc, const float* vfield);

/* Replaces checkpoint.f90:24-47 (the array contents it writes); sim receives
   nplocal/npglobal, sigma_vi, vz_max, dt_* as held by the handle. */
int cube_download_state(cube_handle* h, cube_sim_header* sim, void* xp, void* vp, int32_t* pid,
                        int32_t* rhoc, float* vfield);

/* Replaces update_x -> update_xp (update_particle.f90:5-259) followed by
   buffer_grid and buffer_x (main.f90:57-59). */
int cube_drift(cube_handle* h, float dt_old, float dt, cube_drift_out* out);

/* Replaces ext_pp_force (ext_pp_force.f90:9-160): pair force over fine-cell
   neighbourhoods within +-pp_range, kick and re-quantise in place. */
int cube_pp(cube_handle* h, float a_mid, float dt, cube_pp_out* out);

/* Replaces particle_mesh (pm.f90:2-451) and buffer_v (main.f90:63-64):
   fine and coarse PM forces, kicks, re-quantisation, dt constraints. */
int cube_pm(cube_handle* h, float a_mid, float dt, cube_pm_out* out);

/* Replaces expansion (timestep.f90:136-180): third-order Taylor step of the
   Friedmann equation, real(4) interface, real(8) internals. */
int cube_expansion(float a, float dt, float omega_m, float omega_l, float wde, float* da1, float* da2);

/* Copies the message of the last failing call into buf (NUL-terminated). */
int cube_last_error(cube_handle* h, char* buf, size_t n);

/* Replaces finalize.f90 (plan destruction); frees all device memory. */
int cube_finalize(cube_handle* h);

/* ---- diagnostics with no counterpart in the reference: per-particle force
   probes used by the parity tests and the FP32-vs-FP64 error report.  Each
   writes 3*nplocal doubles in particle order and leaves the state unchanged.
   fine/coarse: sum over the 8 CIC taps of F*w (pm.f90:146-153, 354-361
   without the a_mid*dt/6/pi factor).  pp: the pair sum f_tot of
   ext_pp_force.f90:92-113. */
int cube_probe_pm_forces(cube_handle* h, double* f_fine, double* f_coarse);
int cube_probe_pp_forces(cube_handle* h, double* f_pp);

/* Copies the fine-mesh CIC density of the current positions (nf^3 float,
   x fastest) -- the periodic field pm.f90:76-107 builds tile by tile. */
int cube_probe_density(cube_handle* h, float* rho);

/* Velocity codes of n arguments x = c*v of update_particle.f90:119's atan
   (c = sqrt(pi/2)/(sigma*vrel_boost)), from the device encoder that the
   drift, kick and PP kernels use, for the handle's izipv.  Leaves the state
   unchanged. */
int cube_probe_vencode(cube_handle* h, const double* x, int64_t n, int32_t* codes);

/* Matter power spectrum of the current positions on the fine mesh -- the
   estimator of utilities/cicpower.f90:85-257 with powerspectrum.f90:27-137
   (CIC density, delta = rho/mean - 1, transform / nf^3, sinc^4 window
   correction, linear bins nint(|k|) in grid units, Delta^2 = 4 pi k^3 P).
   nbin must be nint(nf/2*sqrt(3)); for bins 1..nbin (array index b-1) it
   writes the mode count, the sum of |k| and the sum of Delta^2 over modes.
   Overwrites the fine mesh (recomputed by the next cube_pm). */
int cube_power_spectrum(cube_handle* h, int nbin, double* modes, double* ksum, double* delta2sum);

/* No counterpart in the reference.  cube_power_spectrum without any device
   memory allocation, for callers that cannot allocate (code under HIP graph
   capture, tight serving loops, harnesses that forbid hipMalloc).  The
   per-bin partial sums live in a static device array of the library (one
   per GPU, 48 KiB, held until the process ends), so nbin is limited to 2048
   (a 2364^3 mesh); larger nbin is rejected with CUBE_ERR_ARG.  Same
   arguments, results and errors as cube_power_spectrum otherwise.  All
   calls on one GPU share the array: calls that overlap on one GPU race. */
int cube_power_spectrum_noalloc(cube_handle* h, int nbin, double* modes, double* ksum, double* delta2sum);

/* Runtime halo finder -- halofind.f90:18-598 (called from main.f90:88-90
   at the z_halofind redshifts): spherical-overdensity halos of the current
   state, tile by tile as the reference runs it.  Parameters are the
   constants of halofind.f90:27-40 (den_odc = 200 as odc_vir is .false.),
   h0 and box (parameters.f90) and the line-of-sight centre 'cen' (:58);
   a and vsim2phys are the header values sim%a and sim%vsim2phys that
   halofind.f90:438 reads.  nq: pids are q-lattice indices + 1 on an nq^3
   lattice of spacing nf/nq fine cells (the reference's IC numbers the fine
   grid: nq = nf). */
typedef struct cube_halo_params {
  float den_odc, den_peak_cutoff;
  int32_t min_halo_particles, nc_halo_max, newbox_max, search_ratio, refine_ratio, pad;
  int64_t max_halo_np, nq;
  float h0, box, los_centre[3], a, vsim2phys, pad2;
} cube_halo_params;
/* Host buffers for the catalogue: halos[max_halos][42] receives the
   type_halo_info records (halofind.f90:9-15, in the order the reference
   writes them to the 'halo' file), pid_words the 'halo_pid' stream (per
   halo: i_odc, its pids in radius order, 0), hid (optional, nhid int16
   entries) the 'hid' array indexed by pid-1.  On return nhalo, npid_words,
   npeaks (candidate peaks over all tiles) and nprocessed are set. */
typedef struct cube_halo_catalog {
  int64_t max_halos;
  float* halos;
  int64_t max_pid_words;
  int32_t* pid_words;
  int64_t nhid;
  int16_t* hid;
  int64_t nhalo, npid_words, npeaks, nprocessed;
} cube_halo_catalog;
/* Uses the fine mesh as scratch (recomputed by the next cube_pm).  Errors:
   CUBE_ERR_CAPACITY when a catalogue buffer or a search list is too small,
   CUBE_ERR_CHECK for the reference's own stops (search box beyond the tile
   buffer, newbox > newbox_max). */
int cube_halofind(cube_handle* h, const cube_halo_params* p, cube_halo_catalog* out);

/* Seeded Zel'dovich initial conditions written straight into the resident
   state -- the role of the reference's IC tool (util/initial_conditions.f90:
   214-954), which writes CUBE's first checkpoint.  Noise: unit Gaussian per
   lattice site from SplitMix64(SplitMix64(seed)+site) and Box-Muller;
   displacement psi_k = i k/k^2 sqrt(ng^3 P(k)/box^3) noise_k on the
   np_nc*nc lattice; x = q + psi, v = psi*vfac; cells in lattice order.
   Replaces the state; sim_out (optional) receives the header. */
typedef struct cube_ic_params {
  uint64_t seed;
  int32_t np_nc;      /* particles per coarse cell per dimension              */
  int32_t nk;         /* entries of the power-spectrum table (>= 2)           */
  float a;            /* scale factor of the state                            */
  float box;          /* box size, Mpc/h                                      */
  double vfac;        /* velocity per fine-cell displacement, a^2 H(a) f(a)   */
  const double* lnk;  /* ln k, k in h/Mpc, increasing                         */
  const double* lnp;  /* ln P(k), (Mpc/h)^3, at scale factor a                */
  float f_cdm;        /* omega_cdm/omega_m: mass_p_cdm = f_cdm nf^3/np
                         (initial_conditions.f90:943); 0 reads as 1          */
  int32_t reserved;
} cube_ic_params;
int cube_ic_zeldovich(cube_handle* h, const cube_ic_params* ip, cube_sim_header* sim_out);

/* ---- neutrino fluid (hydrodf.f90, hydrogr.f90, hydro3d.f90, hydro1d.f90):
   the dispersive sum of three fluids at the Gauss-Laguerre velocities
   v_n/beta*c, beta = 5966.34 m/eV, on nf/2, nf/4, nf/8 cells per dimension;
   one neutrino mass (parameters.f90:80-83, Nneu = 1). */
typedef struct cube_nu_params {
  float m_ev;          /* neutrino mass, eV (Mneu)                          */
  float f_neu;         /* omega_nu/omega_m (parameters.f90:92)              */
  float z_i_nu;        /* redshift where the fluid starts (parameters.f90:73) */
  float gamma;         /* adiabatic index, 5/3 in main.f90:38               */
  float box, h0, omega_m;  /* for the velocity unit hg_dv (hydrogr.f90:14) */
  int32_t neutrino_ic; /* 1: fluids start at z_i_nu (hydrodf.f90:1,14-16)  */
} cube_nu_params;
/* dfld_setup (hydrodf.f90:54-86) + hg_setup (hydrogr.f90:47-81): uniform
   fluids rho=1, m=0, E=cs2/(gamma-1); normalised density weights.  From
   then on cube_pm adds the fluid density to both PM meshes and the PM
   forces to the fluid (pm.f90:55-74, 157-171, 198-242, 374-422). */
int cube_nu_init(cube_handle* h, const cube_nu_params* p);
/* Replaces the neutrino IC tool (utilities/neutrino_ic.f90:40-181, whose
   output main.f90:25-32 reads through dfld_restart): fluid fields from the
   same seeded linear field as cube_ic_zeldovich with the same ip, filtered
   per fluid by 1/(1+l+l^2) (density) and (1+1.5l+2l^2)/(1+l+l^2)^2
   (velocity), l = k/k_fs, block-averaged onto each fluid mesh; momenta
   rho<v>, energy rho cs2/(g-1) + |m|^2/(2 rho).  Needs cube_nu_init; the
   fluid meshes must not be finer than the particle lattice. */
int cube_nu_ic(cube_handle* h, const cube_ic_params* ip);
/* Geometry of fluid n (0..2): cells per dimension, cell size in fine cells,
   cs^2 in simulation units, density weight. */
int cube_nu_info(cube_handle* h, int n, int32_t* nc, float* dx, float* cs2, float* weight);
/* Fields of fluid n as [5][nc^3] float (rho, mx, my, mz, E), x fastest --
   the physical cells of hg_checkpoint/hg_restart (hydrogr.f90:395-446). */
int cube_nu_set_fluid(cube_handle* h, int n, const float* fld);
int cube_nu_get_fluid(cube_handle* h, int n, float* fld);
/* dfld_evolve (hydrodf.f90:90-110): fluids with a >= a_i swept x,y,z
   (dir > 0) or z,y,x (dir < 0) with step dt; *dt_neu = min(*dt_neu, CFL
   step of each evolved fluid) (hydrogr.f90:104-124). */
int cube_fluid_evolve(cube_handle* h, float dt, int dir, float a, float* dt_neu);

/* ---- host loop in native code: main.f90:48-97 with timestep.f90:8-95 in
   real(4).  cube_clock holds the time-stepping state the reference keeps in
   module variables (a, t, tau, dt, dt_old, da, a_mid, dt_neu,
   cur_checkpoint, checkpoint_step, final_step, z_checkpoint). */
#define CUBE_MAX_CHECKPOINTS 64
typedef struct cube_clock {
  float a, t, tau, dt, dt_old, da, a_mid, dt_neu;
  int32_t cur_checkpoint, n_checkpoint, checkpoint_step, final_step;
  int32_t checkpoint_at;  /* checkpoint index reached by the last step, else 0 */
  int32_t pad;
  int64_t istep;
  float z_checkpoint[CUBE_MAX_CHECKPOINTS];
} cube_clock;
/* initialize.f90:40-53,73: a and tau from z_checkpoint[0], cur_checkpoint = 2;
   1 <= n <= CUBE_MAX_CHECKPOINTS, else CUBE_ERR_ARG. */
int cube_clock_init(cube_clock* c, const float* z_checkpoint, int n);
/* One pass of the main-loop body: timestep, drift (update_x), PP if
   enabled, PM, and at a checkpoint step the closing half drift
   (main.f90:74-75).  After the final checkpoint, c->final_step and
   c->checkpoint_step are both set. */
int cube_step(cube_handle* h, cube_clock* c);
/* How cube_step runs drift, PP and PM (no counterpart in the reference,
   whose loop body calls them one after another; for callers that step
   many small boxes, where host round trips and launch costs dominate):
   0 (default) each phase synchronises as the separate calls do; 1 the
   three are launched back to back, their checks made on flags and their
   results read back after one synchronisation per step; 2 as 1, with the
   launches captured into a hipGraph whose kernel arguments are updated
   each step.  Every mode gives the same bits.  Boxes with the neutrino
   fluid always run mode 0. */
int cube_set_step_mode(cube_handle* h, int mode);
/* hipGraph instantiations and in-place updates made by step mode 2 so far. */
int cube_step_stats(cube_handle* h, int64_t* graph_instantiations, int64_t* graph_updates);
/* Device bytes a box of these parameters holds (plus the Zel'dovich
   generator's temporary mesh when ic_np_nc > 0), as used for admission. */
int64_t cube_footprint(const cube_params* p, int32_t ic_np_nc);

/* ---- ensemble serving (no counterpart in the reference): independent
   boxes of uneven size submitted at any time; each is admitted as soon as
   its footprint fits mem_budget and fewer than max_live boxes run, then
   stepped from its Zel'dovich start to the last z_checkpoint on its own
   stream and host thread, concurrently with the other live boxes. */
typedef struct cube_ensemble cube_ensemble;
typedef struct cube_ens_params {
  int32_t device;
  int32_t max_live;     /* concurrent boxes                                  */
  int64_t mem_budget;   /* device bytes for resident boxes; <= 0: 90% free  */
  int32_t step_mode;    /* cube_set_step_mode for every box                  */
  int32_t reserved;
} cube_ens_params;
typedef struct cube_ens_result {
  int64_t id;
  int32_t status;       /* cube_status of the run                            */
  int32_t steps;
  double t_submit, t_admit, t_done;  /* seconds on the engine clock          */
  int64_t nparticles;
  float a_final;
  int32_t pad;
  uint64_t state_hash;  /* FNV-1a of the final xp, vp, pid                   */
  char message[128];
} cube_ens_result;
int cube_ens_create(const cube_ens_params* p, cube_ensemble** out);
int cube_ens_submit(cube_ensemble* e, const cube_params* p, const float* kern_f_table, const float* kern_c_table,
                    const cube_ic_params* ic, const float* z_checkpoint, int n_z, int64_t* id);
/* Moves up to max_results finished runs into out; *n_out = how many. */
int cube_ens_poll(cube_ensemble* e, cube_ens_result* out, int max_results, int* n_out);
int cube_ens_pending(cube_ensemble* e, int* queued, int* live);
double cube_ens_clock(cube_ensemble* e);
/* Runs everything submitted to completion, then frees the engine. */
int cube_ens_destroy(cube_ensemble* e);

/* Per-kernel timing with HIP events recorded on the handle's stream (no
   counterpart in the reference, whose only timers are the per-phase
   system_clock prints, e.g. pm.f90:45,182-183).  cube_profile_read returns
   the number of slots and fills accumulated milliseconds and launch counts
   per slot since the previous read; cube_profile_slot_name names a slot. */
int cube_profile_enable(cube_handle* h, int on);
int cube_profile_read(cube_handle* h, double* ms, int64_t* count, int nslot);
int cube_profile_slot_name(int slot, char* buf, size_t n);

/* HBM copy-bandwidth probe (no counterpart in the reference; a measuring
   aid for scoring kernels against HBM): times a streaming float4 copy of
   `bytes` bytes, iters times after one warm launch, in four variants
   ({plain, nontemporal} x {8, 32} workgroups per CU) and writes each
   variant's best rate, read plus write bytes per second / 1e9, to
   gbps[0..3].  Allocates and frees 2*bytes of device memory. */
int cube_hbm_copy_probe(int device, int64_t bytes, int iters, double* gbps);

#ifdef __cplusplus
}
#endif
#endif
