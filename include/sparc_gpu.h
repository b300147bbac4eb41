/*
 * sparc_gpu.h -- C ABI of the MI355X (gfx950) time-stepping engine for the
 * SPARC linearized Euler wave solver.
 *
 * The reference has no FFI: its hot path is `call step()` inside TIMESTEPPING
 * (driver.f90:412-424) operating on module globals (initialize.f90:79,123),
 * preceded by mp_Initialize_RHS[_2d] + Initialize_step (driver.f90:220-223).
 * Every entry point below replaces one piece of that in-process seam, so a
 * Fortran host (ISO_C_BINDING) or Python (ctypes) can keep the reference's
 * own set-up and call sequence and hand the per-step work to the GPU.
 *
 * Conventions
 *   - Arrays are Fortran column-major (nx, ny, nz), x fastest, FP64 on the
 *     host side, copied in/out (no host pointer is retained).
 *   - Plane indices (e_rad, o_rad, k) are 1-based like the reference.
 *   - Every call returns SPARC_OK (0) or a negative SPARC_E* code; the text
 *     of the last error of a context is returned by sparc_last_error().
 *   - One context is single-threaded (not re-entrant).  Independent contexts
 *     run on their own HIP streams and may be driven from separate threads.
 */
#ifndef SPARC_GPU_H
#define SPARC_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPARC_ABI_VERSION 1

/* return codes */
#define SPARC_OK 0
#define SPARC_EINVAL (-1)   /* bad argument / unsupported configuration   */
#define SPARC_EHIP (-2)     /* HIP runtime error (text in last_error)      */
#define SPARC_ENOMEM (-3)   /* device allocation failed                    */
#define SPARC_ESTATE (-4)   /* call out of order (e.g. step before setup)  */

/* arithmetic of the state / stepping */
#define SPARC_FP64 0
#define SPARC_FP32 1

/* physics option, numbering of PARSE_OPTIONS (initialize.f90:778-824) */
#define SPARC_OPT_QUIET_PML_3D 2         /* (rho,v,p) form, pml.f90:12-76 */
#define SPARC_OPT_MHD_PML_3D 8           /* (rho,v,p,b) form, pml.f90:80-208 */
#define SPARC_OPT_QUIET_PML_3D_DISPL 4   /* displacement.f90:482-658 */
#define SPARC_OPT_MHD_PML_3D_DISPL 10    /* displacement.f90:663-849 */
#define SPARC_OPT_QUIET_PML_2D_DISPL 17  /* physics2d.f90:948-1105   */
#define SPARC_OPT_MHD_PML_2D_DISPL 18    /* physics2d.f90:724-943    */

/* solve direction, compute_forward / compute_adjoint (initialize.f90:1006-1045) */
#define SPARC_FORWARD 0
#define SPARC_ADJOINT 1

/* state field ids: the dimfive=6 displacement layout (initialize.f90:537-549) */
#define SPARC_XI_X 0
#define SPARC_XI_Y 1
#define SPARC_XI_Z 2
#define SPARC_V_X 3
#define SPARC_V_Y 4
#define SPARC_V_Z 5
#define SPARC_NFIELDS 6
/* option 2 keeps (rho, v, p) in the first five slots (initialize.f90:498-502);
 * it is a forward-only formulation (the kernels need the displacement) */
#define SPARC_RHO 0
#define SPARC_P 4
/* option 8 ((rho, v, p, b) form) adds b in slots 5..7 (initialize.f90:544-551) */
#define SPARC_BX 5
#define SPARC_MAXFIELDS 8

/* which array sparc_get_state / sparc_put_state touch */
#define SPARC_ARR_A 0        /* a(:,:,:,f): the state                          */
#define SPARC_ARR_SCR 1      /* scr(:,:,:,f): last stage RHS ("acc" for f=3..5) */
#define SPARC_ARR_PMLVARS 2  /* pmlvars(:,:,:,f), f=0 psivz, f=1 psip; the MHD
                                options add 2 psidzbx, 3 psidzby,
                                4 psiinductionbx, 5 psiinductionby
                                (initialize.f90:485-565)                        */
#define SPARC_ARR_SCRPML 3   /* scrpml(:,:,:,f)                                 */
#define SPARC_ARR_ALO 4      /* low-order part of a(:,:,:,f) when the FP32 state is
                                compensated (cfg.compensated), as FP64 values;
                                put rounds them to the 16-bit low words under the
                                current high words (clamped to +-32767 steps of
                                ulp/2^16).  get/put of SPARC_ARR_A then
                                move hi+lo as one FP64 value                    */

/* 3-D background fields (sparc_set_background).  n = nx*ny*nz for a full
 * 3-D field or n = nz for a horizontally uniform z-profile. */
#define SPARC_BG_RHO0 0        /* rho0                                        */
#define SPARC_BG_RHOINV 1      /* 1/rho0 (physics.f90:204)                    */
#define SPARC_BG_C2RHO0 2      /* c2*rho0 (physics.f90:205)                   */
#define SPARC_BG_GRADP0_Z 3    /* d p0/dz, PML-decayed (physics.f90:207,271)  */
#define SPARC_BG_GRADRHO0_Z 4  /* d rho0/dz, PML-decayed (physics.f90:206,272)*/
#define SPARC_BG_AZ 5          /* az    = -dampz       (physics.f90:291)      */
#define SPARC_BG_BZPML 6       /* bzpml = -alphaz-dampz (physics.f90:290)      */
#define SPARC_BG_C2 7          /* c2, used by the sound-speed/density kernels */
/* option 10: background field B0 = (box, boy, boz), its curl and the
 * horizontal background gradients (physics.f90:117-190) */
#define SPARC_BG_BX 8
#define SPARC_BG_BY 9
#define SPARC_BG_BZ 10
#define SPARC_BG_CURLBX 11
#define SPARC_BG_CURLBY 12
#define SPARC_BG_CURLBZ 13
#define SPARC_BG_GRADP0_X 14
#define SPARC_BG_GRADP0_Y 15
#define SPARC_BG_GRADRHO0_X 16
#define SPARC_BG_GRADRHO0_Y 17
#define SPARC_BG_COUNT 18

/* z-profiles (sparc_set_profile), length nz */
#define SPARC_PROF_STRETCH 0   /* 1/(dz/dk) (initialize.f90:464-468)          */
#define SPARC_PROF_UNSTRETCH 1 /* dz/dk                                        */
#define SPARC_PROF_G 2         /* gravity, PML-decayed (physics.f90:273)      */
#define SPARC_PROF_Z 3         /* grid z (units of Rsun)                      */
#define SPARC_PROF_SOURCE_DEP 4 /* source_dep(k) = exp(-4e7 (z(k)-z(e_rad))^2)
                                  (physics.f90:109): option 2's source shape */
#define SPARC_PROF_SPONGE 5    /* spongez(k), planes 1..90 zeroed (physics.f90:103-109,
                                  293); option 10 applies it for k >= nz-35     */
#define SPARC_PROF_COUNT 6

/* source kinds (sparc_set_source) */
#define SPARC_SRC_NONE 0
#define SPARC_SRC_SERIES 1   /* mask(x,y)*series(time+1)*delta_width/rho0 at e_rad
                                (physics2d.f90:1064-1068)                        */
#define SPARC_SRC_LAGRANGE 2 /* forcing = 7-point Lagrange of vr(x,y,t) sampled
                                every cadforcing (physics.f90:678-737), injected
                                as forcing*delta_width/rho0 at o_rad
                                (displacement.f90:626, physics2d.f90:1060-1062) */

/* single operators (sparc_apply_op), the reference's derivative/filter
 * subroutines applied to one 3-D field */
#define SPARC_OP_DDX 0       /* ddx (derivatives.f90:28-58): FFT, times stretchx  */
#define SPARC_OP_DDY 1       /* ddy (derivatives.f90:62-170): FFT, times stretchy */
#define SPARC_OP_DDZ 2       /* ddz (derivatives.f90:10-24): dbyd2(ibc)*stretch   */
#define SPARC_OP_ZFILTER 3   /* 11-point z filter, planes 2..nz-1 (kernels.f90:845-881) */
#define SPARC_OP_DAMP 4      /* IFFT2(damping_rates*FFT2(f)) (damping.f90:135-160),
                                planes 2..nz-1                                  */
#define SPARC_OP_LOWPASS_X 5 /* Fermi low-pass along x (physics.f90:796-802)     */
#define SPARC_OP_LOWPASS_Y 6 /* Fermi low-pass along y (physics.f90:810-828)     */

/* kernel classes timed by sparc_profile_read (in the time step each class is
 * one kernel) */
#define SPARC_KC_XLINES 0   /* x-line FFT derivative                            */
#define SPARC_KC_YLINES 1   /* y-line FFT derivative                            */
#define SPARC_KC_ZSOLVE 2   /* compact-FD z solve with the fused momentum and
                               continuity RHS epilogue (and plain stores)      */
#define SPARC_KC_DAMP 3     /* damping: forward x transform pass                */
#define SPARC_KC_UPDATE 4   /* a += dt*scr with the z filter                    */
#define SPARC_KC_OTHER 5    /* sources, PML update, kernel accumulation         */
#define SPARC_KC_ZSOLVE_P 6 /* compact-FD z solve with the fused pressure and
                               density epilogue                                 */
#define SPARC_KC_DAMP_Y 7   /* damping: y transform, rates, inverse y           */
#define SPARC_KC_DAMP_XI 8  /* damping: inverse x transform and RHS update      */
#define SPARC_KC_FILTER 9   /* spectral low-pass along x or y (every 60 s)      */
#define SPARC_KC_COUNT 10

/* kernels returned by sparc_kernel_get (kernels.f90:276-313,530-537) */
#define SPARC_KERN_C2 0       /* sum div(xi_f) div(v_a)                      */
#define SPARC_KERN_RHO 1      /* density-kernel integrand sum                */
#define SPARC_KERN_HESSIAN 2  /* sum |acc_f|^2                               */
#define SPARC_KERN_VX 3       /* flow kernels (cfg.flow_kernels):            */
#define SPARC_KERN_VY 4       /*   sum_j d_i v_f,j * v_a,j, i = x, y, z      */
#define SPARC_KERN_VZ 5       /*   (kernels.f90:317-332)                     */
#define SPARC_KERN_BX 6       /* magnetic kernels of the MHD displacement    */
#define SPARC_KERN_BY 7       /*   options 10/18 (kernels.f90:336-430),      */
#define SPARC_KERN_BZ 8       /*   unscaled sums                             */
#define SPARC_KERN_PSI 9      /* stream-function kernel: y component of      */
                              /*   CURL_KERN(K_Bx, 0, K_Bz), formed from the */
                              /*   sums when read (kernels.f90:508-509)      */
#define SPARC_KERN_COUNT 10

typedef struct sparc_config {
  int32_t abi_version;    /* must be SPARC_ABI_VERSION                          */
  int32_t nx, ny, nz;     /* params.i:25                                        */
  int32_t option;         /* SPARC_OPT_*                                        */
  int32_t precision;      /* SPARC_FP64 / SPARC_FP32                            */
  int32_t npmlbot;        /* params.i:120                                       */
  int32_t npmltop;
  int32_t damp_waves;     /* DAMP_WAVES (params.i:52); option 17 never damps    */
  int32_t freq_filtering; /* floor(60/timestep) (driver.f90:270)                */
  int32_t e_rad;          /* 1-based source plane                               */
  int32_t o_rad;          /* 1-based observation plane                          */
  int32_t mode;           /* SPARC_FORWARD / SPARC_ADJOINT                      */
  int32_t cadforcing_step;/* floor(cadforcing/timestep) (initialize.f90:977)    */
  int32_t forcing_length; /* source samples (initialize.f90:981)                */
  int32_t num_steps;      /* Lagrange end guard (driver.f90:280-281)            */
  int32_t compensated;    /* FP32 only: hold the state as an FP32 word plus a
                             16-bit low word in steps of its ulp/2^16 (40
                             significant bits) and run a += deltat*scr and the
                             z filter in FP64 (the reference state is FP64);
                             0 = plain FP32 state                               */
  int32_t flow_kernels;   /* snapshots also keep v (9 fields) and the kernel sums
                             add the flow kernels SPARC_KERN_VX..VZ            */
  int32_t use_graphs;     /* replay each step's five stages and update as a
                             captured HIP graph (forcing and spectral filter
                             stay eager); 0 = launch every kernel eagerly    */
  int32_t plane_damping;  /* option 4, FP32, 256 x 256 planes: DAMP_VELOCITY
                             in one pass per plane (k_damp2d) instead of the
                             forward x, y and inverse x passes; ignored
                             elsewhere.  0 = the three passes (default)      */
  int32_t reserved_i[4];
  double xlength, ylength;/* box sides, cm (params.i:27)                        */
  double timestep;        /* s (params.i:27)                                    */
  double deltat;          /* timestep*dimc/diml (initialize.f90:472)            */
  double stretchx;        /* rsun/xlength (initialize.f90:223)                  */
  double stretchy;        /* rsun/ylength                                       */
  double delta_width;     /* 2/(z(k+1)-z(k-1)) at the injection plane           */
  double cadforcing;      /* s (params.i:36)                                    */
  double field_scale;     /* FP32 only: state is stored multiplied by this;
                             host copies in/out are unscaled.  1.0 = none.      */
  double reserved_d[7];
} sparc_config;

typedef struct sparc_ctx sparc_ctx;

/* ---- life cycle ------------------------------------------------------- */

/* Allocates the state a/temp_step/scr, PML arrays and work buffers on the
 * current HIP device (Initialize_all allocations, initialize.f90:384-389) and
 * sets the LDDRK coefficients (INITIALIZE_STEP, step.f90:24-41).  The state
 * starts at zero, time at 0. */
int sparc_create(const sparc_config *cfg, sparc_ctx **out);
void sparc_destroy(sparc_ctx *ctx);
const char *sparc_last_error(const sparc_ctx *ctx);
/* Device bytes held by the context (for admission by HBM footprint). */
int sparc_device_bytes(const sparc_ctx *ctx, size_t *bytes);
/* Bytes a context with this configuration would allocate (no allocation). */
int sparc_footprint(const sparc_config *cfg, size_t *bytes);

/* ---- set-up: the arrays mp_Initialize_RHS leaves in module globals ---- */

/* id = SPARC_BG_*; n = nx*ny*nz (3-D) or nz (z-profile, stored compactly). */
int sparc_set_background(sparc_ctx *ctx, int id, const double *host, size_t n);
/* id = SPARC_PROF_*; n = nz. */
int sparc_set_profile(sparc_ctx *ctx, int id, const double *host, int n);
/* damping_rates((nx/2+1), ny) of INIT_DAMPING (damping.f90:38); normx*normy
 * included as in the reference. */
int sparc_set_damping(sparc_ctx *ctx, const double *rates, size_t n);
/* decay_coeff_x(nx/2+1), decay_coeff_y(ny/2+1) (initialize.f90:215-221). */
int sparc_set_filter(sparc_ctx *ctx, const double *decay_x, const double *decay_y);
/* SPARC_SRC_SERIES: mask = vr(:,:,1) (nx*ny), series = fwdsource (nt values,
 *   series[t] used at time t), active while time <= active_until.
 * SPARC_SRC_LAGRANGE: mask = vr(nx,ny,nt) samples (nt = forcing_length+8,
 *   zero padded), series unused; active while time <= active_until. */
int sparc_set_source(sparc_ctx *ctx, int kind, const double *mask, size_t nmask,
                     const double *series, int nt, int active_until);
/* Per-pixel source planes of the MHD displacement options: erad_2d (forward)
 * or orad_2d (adjoint), 1-based, nx*ny values, from the Wilson-depression
 * search of mp_Initialize_RHS (physics.f90:66-90); delta_width_2d(i,j) =
 * 2/(z(k+1)-z(k-1)) is formed from the z profile (set it first).  n = 0
 * returns to the uniform plane e_rad / o_rad. */
int sparc_set_source_planes(sparc_ctx *ctx, const int32_t *k_xy, size_t n);

/* ---- the hot path ------------------------------------------------------ */

/* Advances nsteps full LDDRK steps starting at time index t0: for each
 * i = t0..t0+nsteps-1, time = i; [adjoint: LAGRANGE_INTERP]; STEP (5 RHS
 * stages, a += deltat*scr, pmlvars += deltat*scrpml, FILTER_VARS_3D/_2D).
 * Asynchronous on the context stream. */
int sparc_step(sparc_ctx *ctx, int t0, int nsteps);
/* Blocks until the context's queued work is done. */
int sparc_sync(sparc_ctx *ctx);
/* Zeroes a, scr, pml arrays (a fresh solve in an existing context). */
int sparc_reset(sparc_ctx *ctx);

/* Applies one operator to in (nx*ny*nz) -> out on the device, through the
 * kernels the stepper uses.  For SPARC_OP_DDZ with ibc 2/3/4 the caller's
 * boundary plane(s) are read from out on entry, as dbyd2 expects
 * (dbyd2.f90:128,140).  Synchronous; uses the context's work arrays. */
int sparc_apply_op(sparc_ctx *ctx, int op, int ibc, const double *in, double *out);

/* ---- state in/out (checkpoint + snapshot contract, all_modules.f90:750-784) */
int sparc_get_state(sparc_ctx *ctx, int arr, int field, double *host);
int sparc_put_state(sparc_ctx *ctx, int arr, int field, const double *host);
/* a(:,:,k,field) for one 1-based plane k -> host (nx*ny). */
int sparc_get_plane(sparc_ctx *ctx, int field, int k, double *host);
/* sqrt(sum a(:,:,:,field)^2 / (nx*ny*nz)) (all_modules.f90:176-199). */
int sparc_norm(sparc_ctx *ctx, int field, double *out);
/* Samples a(i,j,k,field) at npts (0-based i,j; 1-based k) points -> host. */
int sparc_sample(sparc_ctx *ctx, int field, int npts, const int32_t *ijk,
                 double *host);
/* sparc_sample without any device allocation, for callers that cannot
 * allocate in their loop (a serving loop sampling receivers every output
 * cadence, code under stream capture).  No counterpart in the reference.
 * The point list and the values are staged in a fixed device array of
 * SPARC_SAMPLE_NOALLOC_MAX points (20 bytes each, 1.25 MiB per GPU, held
 * until the process ends); larger npts returns SPARC_EINVAL.  Every context
 * on a GPU shares that array: calls within one process are serialised by a
 * lock, and the call returns only after the values are on the host. */
#define SPARC_SAMPLE_NOALLOC_MAX 65536
int sparc_sample_noalloc(sparc_ctx *ctx, int field, int npts, const int32_t *ijk,
                         double *host);

/* ---- sensitivity kernels in the time loop (kernels.f90:197-461) -------- */

/* Reserves room for nslots forward snapshots (xi, acc = scr(4..6)) kept in
 * device memory at the working precision. */
int sparc_snapshots_alloc(sparc_ctx *ctx, int nslots);
/* Stores the current forward snapshot (after a step: a(1..3), scr(4..6),
 * WRITE_OUT_PARTIAL_STATE) into slot. */
int sparc_snapshot_store(sparc_ctx *ctx, int slot);
/* Copies the snapshot bank of src into dst (allocating room in dst), ordered
 * against both streams like sparc_kernel_accumulate_from. */
int sparc_snapshots_copy(sparc_ctx *dst, const sparc_ctx *src);
/* Like sparc_kernel_accumulate, with the forward snapshot read from the bank
 * of another context on the same device (src may equal ctx).  Stream-ordered
 * both ways without blocking the host: the read waits for src's queued work,
 * and src's later work (a snapshot store into the same slot) waits for the
 * read. */
int sparc_kernel_accumulate_from(sparc_ctx *ctx, const sparc_ctx *src, int slot);

/* ---- forward-wavefield checkpoints (restart state of driver.f90:489-524,
 * WRITE_OUT_FULL_STATE all_modules.f90:623-694, kept in device memory) ---- */

/* Reserves nslots restart states: a(1..6) (with its low words when the
 * state is compensated) and pmlvars(1..2).  The next step
 * needs nothing else (the first LDDRK stage uses a alone, step.f90:52-56). */
int sparc_checkpoints_alloc(sparc_ctx *ctx, int nslots);
int sparc_checkpoint_save(sparc_ctx *ctx, int slot);
int sparc_checkpoint_load(sparc_ctx *ctx, int slot);

/* Zeroes the FP64 kernel accumulators; sets the gravity profile grav(nz)
 * used by the density kernel (unscaled by the PML, kernels.f90:298-301). */
int sparc_kernel_begin(sparc_ctx *ctx, const double *grav, int nz);
/* Adds one snapshot pair: current adjoint state (v = a(4..6)) with forward
 * slot (kernels.f90:276-313), z-derivatives filtered by FILTER_VAR_Z. */
int sparc_kernel_accumulate(sparc_ctx *ctx, int slot);
/* FP64 accumulator -> host (nx*ny*nz), unscaled sums (no dimen factors). */
int sparc_kernel_get(sparc_ctx *ctx, int which, double *host);

/* ---- instrumentation (hipEvent timers, no reference counterpart beyond the
 * per-step MPI_WTIME prints of driver.f90:416-446) ------------------------ */

/* on != 0: bracket every launch of this context with HIP events on its
 * stream and clear previous records; on == 0: stop recording. */
int sparc_profile_enable(sparc_ctx *ctx, int on);
/* Totals for kernel class cls since sparc_profile_enable(ctx,1): summed event
 * time (ms), algorithmic bytes (the arrays a launch must read and write, not
 * its internal scratch traffic) and the number of launches.  Synchronizes. */
int sparc_profile_read(sparc_ctx *ctx, int cls, double *ms, double *bytes,
                       int64_t *launches);

#ifdef __cplusplus
}
#endif
#endif /* SPARC_GPU_H */
