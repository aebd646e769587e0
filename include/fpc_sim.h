/*
 * fpc_sim.h -- C ABI of the one-time host setup and of the time-step driver (libfpc.so).
 *
 * The setup restates the reference's CPU geometry / assembly / embedding
 * (immersed_geometry.hpp:1-160, immersed_operators.hpp:1-287, 397-447, 737-748,
 * solver_helpers.hpp:51-84, 227-329, main.cu:20-121) directly on the structured
 * grid, without sparse maps, and reproduces the embedded stencils, masks and
 * index sets bit for bit -- including the low bits that the reference's
 * std::unordered_map iteration order gives Ap's diagonal and the l1 row scales.
 *
 * The driver mirrors main.cu:8-424: the same Struct objects and solver settings
 * (through fpc_struct.h) and the same per-step order (main.cu:249-376), with
 * the dense ops run as the fused stages of fpc_dense.h.
 */
#ifndef FPC_SIM_H
#define FPC_SIM_H

#include "fpc_dense.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Params (common.hpp:45-85) plus extensions the production configurations need. */
typedef struct {
  int nx, ny, nz;
  double Lx, Ly, Lz, rho, nu, Um, dt;
  int nsteps, print_every, inlet_ramp_steps, use_zero_initial_guess, nsub_beta;
  double beta_tol, alpha_tol, xcirc, ycirc, Rcirc;
  int vel_maxit, vel_kdim;
  double vel_tol;
  int vel_jacobi_iters;
  int p_maxit, p_kdim;
  double p_tol;
  int p_pc_maxit, p_pc_relax_type, p_pc_rap_type;
  /* extensions (no reference counterpart) */
  int init_profile;          /* 1: u = inlet profile on active u faces at t = 0 */
  double noise_amp;          /* seeded uniform noise in (-a, a) on active faces */
  unsigned long long seed;
  int mg_precision;          /* 64: FP64 GMRES+PFMG; 32: FP32 PFMG inside FP64 flexible GMRES */
  int vel_print, p_print;
  /* output and device (common.hpp:60, 82-84): the driver executable (fpc_main) writes
   * the cell-centred VTK file at steps 1, nsteps and every max(1, vtk_every)-th step
   * when write_vtk is set; fpc_sim_step itself writes no file. */
  int write_vtk, vtk_every, device;
  char vtk_filename[256];
} fpc_params;

void fpc_default_params(fpc_params *par);
/* Reference flag names (common.hpp:87-127, including -write-vtk, -vtk-every, -vtk and
 * -device) plus -init-profile, -noise-amp, -seed, -mg-precision.  Unknown flags are
 * ignored, as in the reference. */
int fpc_parse_args(int argc, char **argv, fpc_params *par);

/* ---- one-time setup ---- */
typedef struct fpc_setup_s *fpc_setup;
int fpc_setup_create(const fpc_params *par, fpc_setup *setup);
/* Named arrays (same names as the reference variables): beta2d alpha_u2d alpha_v2d
 * fluid2d fluid_p active_{u,v,w,qu,qv,qw} id_{u,v,w,p,qu,qv,qw} rhs_u_bc rhs_v_bc rhs_w_bc
 * div_bc scale_{u,v,w} (1-based active vectors) Au_vals Av_vals Aw_vals Ap_vals (AoS boxes)
 * D_rowptr D_colind D_vals G_rowptr G_colind G_vals uin; plus the structured forms
 * scale_{u,v,w}_box rhs_u_bc_box dcoef div_bc_plane. */
long long fpc_setup_len(fpc_setup s, const char *name);
int fpc_setup_elsize(fpc_setup s, const char *name);
int fpc_setup_copy(fpc_setup s, const char *name, void *dst, long long nbytes);
/* Nu Nv Nw Np Nqu Nqv Nqw n_pruned */
long long fpc_setup_int(fpc_setup s, const char *name);
int fpc_setup_destroy(fpc_setup s);

/* ---- time-step driver ---- */
typedef struct fpc_sim_s *fpc_sim;
int fpc_sim_create(const fpc_params *par, fpc_sim *sim);
/* Runs nsteps full steps (main.cu:249-376). */
int fpc_sim_step(fpc_sim sim, int nsteps);
/* One phase of a step: 0 ramp, 1 explicit stage, 2 momentum solves, 3 scatter,
 * 4 divergence, 5 pressure solve, 6 projection, 7 diagnostics. */
int fpc_sim_phase(fpc_sim sim, int phase);
/* Fields: u v w p u_tilde v_tilde w_tilde rhs_u x_u rhs_v x_v rhs_w x_w rhs_p x_p
 * uin_step qout, and every setup array.  fpc_sim_set also takes
 * "inlet_amplitude" (ny*nz doubles, default 1): the inlet velocity and the
 * divergence inlet term become theta * (their reference values); and "step"
 * (one int): the number of steps taken (the next step's ramp follows from it). */
long long fpc_sim_len(fpc_sim sim, const char *name);
int fpc_sim_copy(fpc_sim sim, const char *name, void *dst, long long nbytes);
int fpc_sim_set(fpc_sim sim, const char *name, const void *src, long long nbytes);
/* Device pointer of a field (no copy).  Each simulation runs on a HIP stream of
 * its own (fpc_sim_stream); synchronise with it before using the pointer on
 * another stream. */
int fpc_sim_device_ptr(fpc_sim sim, const char *name, void **ptr);
/* The simulation's HIP stream (a hipStream_t). */
void *fpc_sim_stream(fpc_sim sim);
/* The simulation's operator data for the fused stages of fpc_dense.h (device pointers
 * owned by the simulation; valid until fpc_sim_destroy). */
int fpc_sim_flow_ops(fpc_sim sim, fpc_flow_ops *ops);
/* step ramp div_l2 umax vmax wmax its_u its_v its_w its_p relres_u relres_v relres_w
 * relres_p fx fy cd cl err_u err_v err_w err_p; its_total_c (c = u, v, w, p): iterations
 * summed over the steps run; the setup counts Np Nu Nv Nw Nqu Nqv Nqw
 * n_pruned; zext_m_c and zext_flagged_c (c = u, v, w, p):
 * wall planes (0 = full planes) and flagged points of the hot copy the operator's solve
 * loops read (planned at its first solve); NaN if unknown. */
double fpc_sim_scalar(fpc_sim sim, const char *name);
/* Per-stage wall time (s) accumulated over steps: t_build t_mom t_p t_proj t_diag */
double fpc_sim_timer(fpc_sim sim, const char *name);
int fpc_sim_destroy(fpc_sim sim);

/* ---- on-disk formats ----
 * The reference driver's cell-centred ASCII VTK file (write_vtk_cellcentered,
 * immersed_operators.hpp:1055-1076), byte for byte: host arrays of the full staggered
 * fields, fluid mask of the cells. */
int fpc_write_vtk_cellcentered(const char *filename, const double *u, const double *v, const double *w,
                               const double *p, const unsigned char *fluid_p, int nx, int ny, int nz, double Lx,
                               double Ly, double Lz);
/* base with "_%06d" (step) before its extension (vtk_filename_with_step,
 * solver_helpers.hpp:470-482); FPC_ERROR_ARG if out (outlen bytes) is too short. */
int fpc_vtk_filename_with_step(const char *base, int step, char *out, int outlen);
/* Raw dump of the staggered fields (no reference counterpart): a 40-byte header
 * ("FPCFLD01", int32 nx ny nz step, int64 number of doubles, 4 zero bytes), then u, v, w
 * and p as little-endian float64, full staggered layouts, x fastest.  Host arrays. */
int fpc_write_fields(const char *path, int step, int nx, int ny, int nz, const double *u, const double *v,
                     const double *w, const double *p);
/* The same two files from a simulation's current fields. */
int fpc_sim_write_vtk(fpc_sim sim, const char *filename);
int fpc_sim_write_fields(fpc_sim sim, const char *path);

/* ---- ensemble of independent channels (no counterpart in the reference, which
 * runs one channel per process) ----
 * n instances with their own parameters (uneven grids, cylinders, Reynolds
 * numbers).  nthreads >= 1: each instance on its own HIP stream, created and
 * stepped concurrently by a pool of nthreads host worker threads (largest
 * instance first).  nthreads = 0: batched on one stream (at most 256
 * instances): the momentum solves of all instances and their pressure solves
 * each advance with one launch per solver kernel and V-cycle level over every
 * instance; needs the reference solver settings (Jacobi with one sweep, the
 * FP64 PFMG-style V-cycle with one sweep each way and one cycle).  Either way
 * every instance computes exactly what it computes alone.  Kernel timing
 * (fpc_ProfileEnable) must be off while an ensemble steps. */
typedef struct fpc_ensemble_s *fpc_ensemble;
int fpc_ensemble_create(const fpc_params *pars, int n, int nthreads, fpc_ensemble *ens);
/* Advances every instance by nsteps (a solver's maximum-iterations flag is not an error). */
int fpc_ensemble_step(fpc_ensemble ens, int nsteps);
int fpc_ensemble_size(fpc_ensemble ens);
/* Batched ensembles: the launch configuration chosen by the library's autotuner
 * at creation (it times one V-cycle of every pressure system under each
 * candidate): z-chunk length of the z-marching kernels and the smallest box
 * (points) that runs them.  The choice changes speed only, never results. */
int fpc_ensemble_tiles(fpc_ensemble ens, int *kz, long long *tile_min);
fpc_sim fpc_ensemble_member(fpc_ensemble ens, int i);
/* Packed offset table of field u, v, w or p over the instances: instance i owns
 * elements offsets[i] .. offsets[i+1]-1 of a packed array (n+1 entries). */
int fpc_ensemble_offsets(fpc_ensemble ens, const char *field, long long *offsets);
/* The field of every instance copied into one packed host array (offset table layout). */
int fpc_ensemble_pack(fpc_ensemble ens, const char *field, double *dst, long long nbytes);
int fpc_ensemble_destroy(fpc_ensemble ens);

/* ---- adjoint (no counterpart in the reference) ----
 * Discrete adjoint of the time step of a simulation, for gradients of the
 * time-averaged drag coefficient with respect to the inlet amplitude field
 * theta (fpc_sim_set(sim, "inlet_amplitude", theta[ny*nz]): the inlet is
 * theta * Um * profile) and the viscosity nu.  The transposed systems use the
 * simulation's solver settings. */
typedef struct fpc_adjoint_s *fpc_adjoint;
int fpc_adjoint_create(fpc_sim sim, fpc_adjoint *adj);
/* Vector-Jacobian product of step `step` (1-based) taken from the state
 * (u, v, w, p) (device arrays of the full staggered sizes): on entry (au, av,
 * aw, ap) hold the adjoint of the state after the step, on return that of the
 * state before it; theta_bar (device, ny*nz) and *nu_bar (host) are
 * accumulated.  Leaves the simulation's fields at the recomputed step's start. */
int fpc_adjoint_step_vjp(fpc_adjoint adj, int step, const double *u, const double *v, const double *w,
                         const double *p, double *au, double *av, double *aw, double *ap, double *theta_bar,
                         double *nu_bar);
/* From the simulation's current state: runs nsteps steps with
 * J = mean of c_D over steps avg_from..nsteps (1-based, relative), then its
 * gradient by reverse steps with the state checkpointed every `every` steps and
 * each segment recomputed.  grad_theta: host, ny*nz.  The simulation ends at the
 * state after the nsteps steps. */
int fpc_adjoint_drag_gradient(fpc_adjoint adj, int nsteps, int avg_from, int every, double *J, double *grad_theta,
                              double *grad_nu);
int fpc_adjoint_destroy(fpc_adjoint adj);

#ifdef __cplusplus
}
#endif
#endif /* FPC_SIM_H */
