/*
 * diablo_gpu.h -- C-ABI of the MI355X implementation of DIABLO 3's channel
 * RK3/Crank-Nicolson substep (rk_chan_1) and of the CPU oracle that mirrors it.
 *
 * Reference interface being replaced (reference tree, Jamie-Hilditch/diablo3_test):
 *   - for/channel.f90:10-917    subroutine rk_chan_1 (no arguments; module globals)
 *   - for/diablo.f90:98-111     the driver loop that calls it 3x per step and then
 *                               advances `time` by delta_t
 *   - for/channel.f90:946-1023  subroutine rem_div_chan
 *   - for/boundary.f90:824-838  subroutine apply_BC_vel_mpi_post
 *   - for/advance.f90:17-44     subroutine pre_first_step (post BCs + rem_div)
 * The reference passes everything through module globals; here the globals that
 * the substep reads become `dgpu_params` + the y-grid, and the state arrays
 * (flow.f90:20-25, 50-51) are copied in and out through dgpu_set_field /
 * dgpu_get_field in the reference's own memory layout:
 *
 *   spectral field  complex128, Fortran order (0:Nx/2, 0:Nz+1, 0:Ny+1)
 *                   == (Nx+2)*(Nz+2)*(Ny+2) doubles (fft.f90:72-97, alloc_array3D);
 *                   only i<=Nx/3, k<=2*(Nz/3) carry data (fft.f90:300-313)
 *   physical field  float64, Fortran order (0:Nx+1, 0:Nz+1, 0:Ny+1)   (nu_t, kappa_t)
 *   scalars         the field array repeated n_th times (alloc_array4D, fft.f90:101-129)
 *
 * All entry points return 0 on success and a negative DGPU_E* code on error;
 * dgpu_last_error() returns the message of the last failure on that handle.
 * No C++ exception crosses this boundary.  Calls on one handle are not
 * re-entrant; separate handles are independent.  One handle owns all of its
 * device memory and one HIP stream; host buffers are borrowed for the duration
 * of a call only.
 *
 * The CPU oracle (oracle/liboracle.so) exports the same functions with the
 * prefix `orc_` instead of `dgpu_`; it is test infrastructure only.
 */
#ifndef DIABLO_GPU_H
#define DIABLO_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The most passive scalars a handle takes (the reference fixes N_th at
   compile time, make_headers:17-21); the per-scalar arrays of dgpu_params
   hold this many entries. */
#define DGPU_MAX_TH 16

/* Error codes. */
#define DGPU_OK 0
#define DGPU_EINVAL (-1)   /* bad argument (size, field id, null pointer, ...) */
#define DGPU_EHIP (-2)     /* a HIP runtime call failed */
#define DGPU_ENOMEM (-3)   /* device allocation failed */
#define DGPU_EUNSUP (-4)   /* configuration not supported by this build */
#define DGPU_ESTATE (-5)   /* non-finite value detected in the state */

/* Precision modes for dgpu_create. */
#define DGPU_FP64 0  /* everything in FP64 (parity path) */
#define DGPU_MIXED 1 /* FFTs, nonlinear products and LES in FP32; solves,
                        mean-pressure-gradient forcing, projection and the
                        state itself in FP64 */

/* Field ids for dgpu_set_field / dgpu_get_field. */
#define DGPU_CU1 0
#define DGPU_CU2 1
#define DGPU_CU3 2
#define DGPU_CP 3
#define DGPU_CTH 4   /* n_th scalars, one after another */
#define DGPU_CF1 5   /* RK3 carry of the explicit terms (cf1..3, cfth) */
#define DGPU_CF2 6
#define DGPU_CF3 7
#define DGPU_CFTH 8
#define DGPU_NU_T 9      /* physical, eddy viscosity at GY points */
#define DGPU_KAPPA_T 10  /* physical, eddy diffusivity, n_th of them */
#define DGPU_NFIELDS 11

/*
 * Run parameters: the module globals of for/parameters.f90:23-82 that
 * rk_chan_1 reads, with the derived values of parameters.f90:101-103 and
 * set_parameters (parameters.f90:241-281) already applied by the caller.
 * Boundary-condition integers: 0 = Dirichlet, 1 = Neumann (channel.f90:1545-1876).
 * v_bc_* is the wall-normal velocity u2, w_bc_* the spanwise u3 (note the
 * TOML V/W swap at parameters.f90:373-382: TOML "V_*" feeds w_bc_*).
 */
typedef struct dgpu_params {
  int32_t nx, ny, nz, n_th;   /* Nx, Ny (wall-normal points), Nz: Nx a power of two in
                                 [16, 4096], Nz any length in [16, 2048] */
  double lx, lz;              /* periodic box lengths */
  double re;                  /* nu = 1/re (parameters.f90:101) */
  double nu_v_scale;          /* vertical/horizontal viscosity ratio */
  double beta;                /* horizontal Laplacian exponent (1 = ordinary viscosity) */
  double ro_inv, delta;       /* f-plane Coriolis 1/Ro and aspect ratio */
  double grav_x, grav_y, grav_z;
  double delta_t;             /* fixed time step (h_bar = delta_t*(8,2,5)/15) */
  int32_t f_type;             /* 1 const px0, 2 oscillating px0, 4 oscillating top stress */
  int32_t les_model_type;     /* 1 Smagorinsky, 4 AMD-Rozema, 5 AMD-Verstappen */
  double px0, omega0, amp_omega0, force_start;
  int32_t u_bc_ymin, v_bc_ymin, w_bc_ymin;
  int32_t u_bc_ymax, v_bc_ymax, w_bc_ymax;
  double u_bc_ymin_c1, v_bc_ymin_c1, w_bc_ymin_c1;
  double u_bc_ymax_c1, v_bc_ymax_c1, w_bc_ymax_c1;
  double pr[DGPU_MAX_TH];
  int32_t th_bc_ymin[DGPU_MAX_TH];
  int32_t th_bc_ymax[DGPU_MAX_TH];
  double th_bc_ymin_c1[DGPU_MAX_TH];
  double th_bc_ymax_c1[DGPU_MAX_TH];
  double dthdx[DGPU_MAX_TH], dthdz[DGPU_MAX_TH]; /* background scalar gradients */
  int32_t use_les;
  int32_t create_new_flow;    /* LES gate: LES only if !create_new_flow || time_step>100 */
  /* flavor == 'Front' (parameters.f90:25): with ro_inv != 0 the substep adds
     the advection by the thermal wind of the background gradients dthdx,
     dthdz, a wind that varies linearly about y = ly/2
     (user_rhs_chan_fourier, forcing.f90:50-100). */
  int32_t flavor_front;
  double ly;                  /* LY, the height of the domain (parameters.f90:31) */
} dgpu_params;

typedef struct dgpu_handle dgpu_handle;

/*
 * Create a solver.  gy holds the Ny+1 wall-normal grid points GY(1..Ny+1)
 * exactly as stored in grid.h5:/grids/y (domain.f90:429-483); GYF, DY, DYF and
 * their ghost values are derived as in domain.f90:230-254, 501-527.
 * Replaces init_mpi/create_grid_chan/init_chan_mpi/init_flow/init_fft
 * (domain.f90:55-337, flow.f90:58-88, fft.f90:133-241) for one GPU.
 */
int dgpu_create(const dgpu_params* p, const double* gy, int precision, dgpu_handle** out);
int dgpu_destroy(dgpu_handle* h);
const char* dgpu_last_error(const dgpu_handle* h);

/* Copy one field between a host buffer in the reference layout and the device. */
int dgpu_set_field(dgpu_handle* h, int field, const double* host);
int dgpu_get_field(dgpu_handle* h, int field, double* host);

/*
 * Physical-space copies of a spectral state field (cu1..3, cp, cth) in the
 * physical reference layout (float64, Fortran order (0:Nx+1, 0:Nz+1, 0:Ny+1),
 * n_th copies for cth).  dgpu_get_physical is `call fft_xz_to_physical(cu, u)`
 * on a copy (the field itself is unchanged); dgpu_set_physical is
 * `call fft_xz_to_fourier(u, cu)` on all planes 0..Ny+1, which also drops the
 * modes outside the dealiased block.  These are the transforms around the
 * flow-file I/O of WriteHDF5/ReadHDF5 (phdf5.f90:1076-1082, 1190-1195,
 * 1431-1446); the file format itself is handled by the host library.
 */
int dgpu_get_physical(dgpu_handle* h, int field, double* host);
int dgpu_set_physical(dgpu_handle* h, int field, const double* host);

/*
 * The same exchange with a buffer in device memory (for the torch side,
 * which passes tensor.data_ptr() and its current stream): dev holds the
 * field in the reference layout above (dgpu_field_doubles(h, field)
 * doubles), and the layout conversion runs on `stream` (a hipStream_t; NULL
 * is the null stream).  The copy is ordered after the work already queued on
 * the handle's stream, and the handle's later work after the copy, through
 * events; the host does not wait.  dgpu_get_field_device writes the whole
 * reference-layout buffer (zeros outside the dealiased block).
 */
int dgpu_set_field_device(dgpu_handle* h, int field, const double* dev, void* stream);
int dgpu_get_field_device(dgpu_handle* h, int field, double* dev, void* stream);

/* apply_BC_vel_mpi_post (boundary.f90:299-498, 824-838) on cu1..3. */
int dgpu_apply_bc_vel_post(dgpu_handle* h, double time);
/* rem_div_chan (channel.f90:946-1023): project cu1..3 onto the discrete
   divergence-free space.  Does not touch cp. */
int dgpu_rem_div(dgpu_handle* h);

/* One RK3 substep: `call rk_chan_1` with rk_step in 1..3 (channel.f90:10-917). */
int dgpu_rk_substep(dgpu_handle* h, int rk_step, double time, int time_step);
/* nsteps full steps of the driver loop (diablo.f90:98-108): time_step += 1,
   three substeps, time += delta_t.  *time and *time_step are updated. */
int dgpu_step(dgpu_handle* h, int nsteps, double* time, int32_t* time_step);

/*
 * poisson_p_chan (channel.f90:1026-1234): cp on planes 1..Ny from the
 * velocity (the Poisson equation for the pressure with right-hand side
 * (du_i/dx_j)(du_j/dx_i) minus the buoyancy term Ri(n) d(th_n)/dy) and the
 * walls' Neumann rows.  ri holds Ri(1..n_th) (parameters.f90:73; may be NULL
 * when n_th = 0).  pre_first_step (advance.f90:17-44) calls it for start
 * files without P.  The reference computes it in the RK carry arrays
 * cf1..cf3, which rk_step 1 never reads; here the carry is left unchanged.
 */
int dgpu_poisson_p(dgpu_handle* h, const double* ri);

/*
 * Variable time step: courant (advance.f90:47-158).
 *
 * dgpu_set_courant turns variable_dt on (c != NULL) or off (c == NULL).
 * While it is on, substep rk_step 3 with time_step % update_dt == 0 runs
 * courant on the velocity just after the Crank-Nicolson solves
 * (channel.f90:882-888): the CFL limit over the grid, the inertial and
 * buoyancy periods, max_dt, and the scheduling towards the next output time,
 * exactly as courant sets delta_t.  The new delta_t (and h_bar, and the CN
 * factor tables) take effect from the next substep; dgpu_step advances time
 * by the new delta_t after the step, as diablo.f90:108 does.
 *
 * dgpu_courant runs courant once on the current state, as pre_first_step
 * does (advance.f90:36-38).  There the reference's u1..u3 share memory with
 * cu1..cu3 (fft.f90:72-97) and hold the spectral coefficients, so courant
 * reads those coefficients as a real array; that is reproduced.
 */
typedef struct dgpu_courant_params {
  double cfl;              /* CFL */
  double max_dt;           /* max_dt: the input DELTA_T (parameters.f90:103) */
  double ly;               /* LY (Front flavour) */
  double time_limit;       /* TIME_LIMIT */
  double save_stats_time;  /* next statistics output time */
  double save_flow_time;   /* next flow output time */
  int32_t update_dt;       /* UPDATE_DT >= 1 */
  int32_t front;           /* 1 when flavor == 'Front' (thermal-wind CFL) */
} dgpu_courant_params;
int dgpu_set_courant(dgpu_handle* h, const dgpu_courant_params* c);
int dgpu_courant(dgpu_handle* h, double time);
/* The time step in use, and a new one (h_bar and the CN factor tables are
   rebuilt; delta_t > 0). */
int dgpu_get_delta_t(dgpu_handle* h, double* delta_t);
int dgpu_set_delta_t(dgpu_handle* h, double delta_t);

/* Wait for all work queued on the handle's stream. */
int dgpu_synchronize(dgpu_handle* h);

/*
 * Finiteness check of the state (no counterpart in the reference, whose run
 * goes on with a non-finite field until its statistics or a flow save show
 * it).  Scans cu1..3, cp, cth, the RK carry cf1..3, cfth and, with LES,
 * nu_t and kappa_t for NaN and Inf on the device; returns DGPU_OK, or
 * DGPU_ESTATE with the offending fields named in dgpu_last_error.  Waits for
 * the handle's stream.
 */
int dgpu_check_state(dgpu_handle* h);

/*
 * dgpu_check_state without any device allocation (no counterpart in the
 * reference: it is for callers that cannot allocate device memory, such as
 * a tight serving loop or a harness that forbids hipMalloc).  Same scan and
 * results as dgpu_check_state, whose per-block flags it keeps in a static
 * device array of 29 KiB (128 blocks x 58 arrays of ints) instead of a
 * buffer the handle allocates on first use; scans of more arrays are
 * rejected with DGPU_EINVAL.  Each GPU holds one such array until the
 * process ends, shared by every call on that GPU: calls that overlap on one
 * GPU race.  Calls no hipMalloc or hipFree.
 */
int dgpu_check_state_noalloc(dgpu_handle* h);

/*
 * Adjoint of the FP64 substep, DNS and LES (no counterpart in the reference,
 * which has no differentiation; it is the added "differentiation" form of
 * the substep for gradient-based studies).  The handle holds an adjoint
 * state with the spectral fields of the forward state (ids
 * DGPU_CU1..DGPU_CFTH, reference layout): for a real loss L of a complex
 * field c it holds dL/dRe c + i dL/dIm c.
 *
 * dgpu_adj_rk_substep: with the forward state equal to the INPUT of substep
 * rk_step (as rk_chan_1 would receive it) and the adjoint state equal to the
 * gradient with respect to that substep's OUTPUT, replace the adjoint state
 * by the gradient with respect to the input.  The forward state is left
 * unchanged.  Without LES (or with its gate closed) the substep is affine in
 * cp and the RK carry, so only cu1..3 and cth of the input are read; with
 * the LES terms the adjoint reruns the forward substep and reads the whole
 * input state.  Returns DGPU_EUNSUP for the mixed form and while
 * variable_dt is on.
 */
int dgpu_adj_set_field(dgpu_handle* h, int field, const double* host);
int dgpu_adj_get_field(dgpu_handle* h, int field, double* host);
int dgpu_adj_zero(dgpu_handle* h);
/* The adjoint state through device buffers, as dgpu_set_field_device. */
int dgpu_adj_set_field_device(dgpu_handle* h, int field, const double* dev, void* stream);
int dgpu_adj_get_field_device(dgpu_handle* h, int field, double* dev, void* stream);
int dgpu_adj_rk_substep(dgpu_handle* h, int rk_step, double time, int time_step);
/* Adds w[j] (j = 0..Ny+1) to the real part of the (0,0) mode of the adjoint
   of field (cu1, cu2, cu3 or cp): the gradient of a loss of the mean profile,
   such as the wall shear stress of statistics.f90:143-146. */
int dgpu_adj_add_mean(dgpu_handle* h, int field, const double* w);
/* Real part of the (0,0) mode of a forward field on planes 0..Ny+1 (the mean
   profile ume of statistics.f90:728), Ny+2 doubles. */
int dgpu_mean_profile(dgpu_handle* h, int field, double* out);
/* Checkpoints of the forward state in device slot `slot` (0..4096,
   allocated on first save), for recomputation in adjoint runs.  mode 1 saves
   the whole state (cu1..3, cp, cth, cf1..3, cfth and the f_type 4 wall
   value), mode 2 only what dgpu_adj_rk_substep reads without LES (cu1..3,
   cth), mode 0 restores what the slot holds. */
int dgpu_checkpoint(dgpu_handle* h, int slot, int mode);

/*
 * Per-kernel-class timing (no counterpart in the reference, whose only timer
 * is the wall clock of control.f90:16-76).  While enabled, every kernel
 * launch is bracketed by HIP events on the handle's stream; dgpu_timing_read
 * returns, per class, the summed kernel time (ms), the summed algorithmic HBM
 * bytes of those launches and the launch count.  Enabling resets the sums.
 */
#define DGPU_K_XFWD 0     /* x-direction forward transform pass */
#define DGPU_K_XINV 1     /* x-direction inverse transform pass */
#define DGPU_K_ZFWD 2     /* z-direction forward transform pass */
#define DGPU_K_ZINV 3     /* z-direction inverse transform pass */
#define DGPU_K_XPROD 4    /* nonlinear products + x transform */
#define DGPU_K_RHS 5      /* z transform of products + explicit RK terms */
#define DGPU_K_THOMAS 6   /* Crank-Nicolson tridiagonal solve */
#define DGPU_K_REMDIV 7   /* pressure Poisson solve + projection */
#define DGPU_K_BCPOST 8   /* wall boundary values */
#define DGPU_K_LES 9      /* subgrid model pointwise kernels and forcing */
#define DGPU_K_ADJOINT 10 /* adjoint-only kernels (projection, CN post, RHS, products) */
#define DGPU_K_COPY 11    /* device copies of the state (checkpoints, LES pre-BC wall planes) */
#define DGPU_K_INIT 12    /* Poisson right-hand side and solve, CFL reductions */
#define DGPU_NCLASS 13
int dgpu_timing_enable(dgpu_handle* h, int enable);
int dgpu_timing_read(dgpu_handle* h, double* ms /*[DGPU_NCLASS]*/, double* bytes /*[DGPU_NCLASS]*/,
                     int64_t* count /*[DGPU_NCLASS]*/);
const char* dgpu_kernel_class_name(int cls);
/* The part of the bytes dgpu_timing_read reports that is x-spectra traffic:
   the intermediate of the two-pass 2-D transforms, written by an x pass and
   read by a z pass (or the reverse).  A one-pass transform moves none of it,
   so bytes - tbytes is the class's traffic under that model. */
int dgpu_timing_read_intermediate(dgpu_handle* h, double* tbytes /*[DGPU_NCLASS]*/);

/* Sizes of the reference-layout buffers (in doubles) for a field id. */
int64_t dgpu_field_doubles(const dgpu_handle* h, int field);

#ifdef __cplusplus
}
#endif

#endif /* DIABLO_GPU_H */
