/*
 * fpc_oracle.h -- C interface of the CPU oracle (test infrastructure only).
 *
 * The oracle is a plain C++ restatement of the reference's per-timestep path
 * (JD63021/ibm-3d-fpc: main.cu:20-376, gpu_dense_ops.hpp, immersed_*.hpp,
 * solver_helpers.hpp) and of the HYPRE Struct solvers it calls.  Only tests,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, always as
 * the checker; the product library never links or calls it.
 */
#ifndef FPC_ORACLE_H
#define FPC_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors the reference Params (common.hpp:45-85) plus the extensions the
 * production configuration needs (initial profile + seeded noise). */
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
  int init_profile;
  double noise_amp;
  unsigned long long seed;
} orc_params;

/* Host threads of the solver and stage loops (default 1: the sequential sums). */
void orc_set_threads(int n);
/* Largest coarsest MG level (0, the default: min(1152, max(64, N0/2048))). */
void orc_set_coarse_size(int n);
void orc_default_params(orc_params *par);

/* One-time setup (geometry, masks, ids, assembly, scaling, embedding). */
void *orc_setup_create(const orc_params *par);
void orc_setup_destroy(void *h);
/* Number of elements of a named setup array, or -1 if unknown. */
long long orc_setup_len(void *h, const char *name);
/* Element size in bytes of a named array (1 = uint8, 4 = int32, 8 = double). */
int orc_setup_elsize(void *h, const char *name);
int orc_setup_copy(void *h, const char *name, void *dst, long long nbytes);
/* Named integer scalars: Nu Nv Nw Np Nqu Nqv Nqw n_pruned */
long long orc_setup_int(void *h, const char *name);

/* y = A x for a 7-point box matrix given as AoS values (7 per point). */
void orc_struct_matvec(int nx, int ny, int nz, const double *vals, const double *x, double *y);

/* CSR y[row+1] = sum vals*x[col] (reference k_spmv_csr_1based). */
void orc_csr_spmv_1based(int nrows, const int *rowptr, const int *colind, const double *vals,
                         const double *x, double *y);

/* GMRES(kdim) with HYPRE's control flow.  precond: 0 none, 1 Jacobi, 2 PFMG-style MG.
 * Returns 0, or 256 when maxit was reached without convergence. */
int orc_gmres(int nx, int ny, int nz, const double *vals, const double *b, double *x,
              double tol, int maxit, int kdim, int precond, int pc_iters,
              int relax_type, int rap_type, int *its, double *relres);

/* One application of the preconditioner (zero initial guess): x = M^-1 b. */
int orc_precond_apply(int nx, int ny, int nz, const double *vals, const double *b, double *x,
                      int precond, int pc_iters, int relax_type, int rap_type);

/* MG hierarchy description: fills dims[3*l..], cdir[l], omega[l]; returns number of levels. */
int orc_mg_levels(int nx, int ny, int nz, const double *vals, int rap_type, int max_levels,
                  int *dims, int *cdir, double *omega);

/* Level l's operator of that hierarchy as AoS (7 per point). */
int orc_mg_level_matrix(int nx, int ny, int nz, const double *vals, int relax_type, int level, double *out);

/* Whole simulation (setup + time loop). */
void *orc_sim_create(const orc_params *par);
void orc_sim_destroy(void *s);
/* Runs the phases of one full step, nsteps times. */
int orc_sim_step(void *s, int nsteps);
/* Individual phases of one step (main.cu:249-363), for stage-level parity:
 *  0 ramp/inlet update for step+1 (sets current step), 1 explicit stage,
 *  2 momentum solves, 3 scatter + BC cleanup, 4 divergence + pressure rhs,
 *  5 pressure solve, 6 projection update, 7 diagnostics + forces. */
int orc_sim_phase(void *s, int phase);
long long orc_sim_len(void *s, const char *name);
int orc_sim_copy(void *s, const char *name, void *dst, long long nbytes);
int orc_sim_set(void *s, const char *name, const void *src, long long nbytes);
/* Named doubles: ramp div_l2 umax vmax wmax relres_u relres_v relres_w relres_p
 * cd cl fx fy; named ints via the same call cast: its_u its_v its_w its_p step */
double orc_sim_scalar(void *s, const char *name);

/* The reference's cell-centred VTK writer (immersed_operators.hpp:1055-1076), fprintf as written. */
int orc_write_vtk_cellcentered(const char *filename, const double *u, const double *v, const double *w,
                               const double *p, const unsigned char *fluid_p, int nx, int ny, int nz, double Lx,
                               double Ly, double Lz);

#ifdef __cplusplus
}
#endif
#endif
