/*
 * ocfd_hip.h -- C ABI of the MI355X (gfx950) right-hand-side + RK3 path of the
 * OpenCFD-SC compressible Navier-Stokes solver.
 *
 * The reference solver is Fortran 90 whose hot-path routines take no dummy
 * arguments and work on module globals (OCFD_parameters.f90:70-108).  This
 * library keeps the same data in device memory and exposes one entry point per
 * reference routine, so the Fortran driver can replace the routine bodies with
 * a single call each (iso_c_binding), plus whole-step entry points for callers
 * that keep the state device-resident.
 *
 * Conventions
 *  - Every entry point returns int: 0 = OK, < 0 = configuration / HIP error
 *    (message available from ocfd_last_error), > 0 = a numerical event
 *    (e.g. the number of cells with T <= 0 found by ocfd_comput_duvwt).
 *  - Host arrays use the reference's Fortran layout: halo'd fields are
 *    X(1-LAP:nx+LAP, 1-LAP:ny+LAP, 1-LAP:nz+LAP), conserved-variable arrays are
 *    f(nx,ny,nz,5); column-major, i fastest.  The library never keeps or frees
 *    a host pointer; it copies in/out synchronously.
 *  - One handle = one block on one GPU; calls on one handle are serialised on
 *    the handle's HIP stream and are not thread-safe.
 */
#ifndef OCFD_HIP_H
#define OCFD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Scheme / splitting / BC codes: same values as OCFD_constants
 * (OCFD_parameters.f90:23-31). */
enum {
    OCFD_SCHEME_WENO5 = 1,
    OCFD_SCHEME_WENO7 = 2,
    OCFD_SCHEME_OMP6 = 3,
    OCFD_SCHEME_UD7L = 4,
    OCFD_SCHEME_CD6 = 5,
    OCFD_SCHEME_CD8 = 6,
    OCFD_SCHEME_HYBRID = 10
};
enum { OCFD_SPLIT_SW = 1, OCFD_SPLIT_LLF = 2 };
enum { OCFD_BC_NONE = 0, OCFD_BC_BLUNT2D = 1, OCFD_BC_BOUNDARYLAYER = 2, OCFD_BC_SWEPTCORNER = 3, OCFD_BC_USER_DEF = 99 };
/* Arithmetic of the right-hand side.  OCFD_PREC_MIXED forms the flux
   splitting, the reconstructions of the sweeps and the viscous derivatives,
   stresses and fluxes and accumulates du in FP32, and stores the primitives
   d, u, v, w, T (derived values, recomputed from f every stage in FP64) as
   FP32; the conserved variables f, fn, the RK3 stage combination, metrics and
   Jacobian stay FP64, and so do the characteristic basis and projections of
   the faces the hybrid sensor flags (their splitting and WENO reconstruction
   are FP32).  The transfers of du
   and of the primitives (upload, download, flow-data files) convert on the
   host; an upload of the primitives forms f from the FP64 values given. */
enum { OCFD_PREC_FP64 = 0, OCFD_PREC_MIXED = 1 };
/* How grid metrics are held on the device. */
enum { OCFD_METRIC_FIELD = 0, OCFD_METRIC_UNIFORM = 1 };

/*
 * Run parameters: the namelist &control_opencfd after read_parameter
 * (OCFD_readpara.f90:3-200), for a single block (npx0=npy0=npz0=1).
 * Derived quantities (hx, Cv, Cp, stencil widths, UD7L coefficients,
 * boundary-scheme flags) are computed by the library exactly as
 * set_parameters (OCFD_readpara.f90:261-341) does.
 */
typedef struct ocfd_config {
    int32_t nx, ny, nz;            /* nx_global, ny_global, nz_global          */
    int32_t lap;                   /* halo width LAP; must be 4                 */
    int32_t iperiodic[3];          /* Iperiodic_X/Y/Z (1 = periodic)           */
    int32_t scheme_invis;          /* OCFD_SCHEME_WENO5/WENO7/OMP6/UD7L/HYBRID  */
    int32_t scheme_vis;            /* OCFD_SCHEME_CD6 / OCFD_SCHEME_CD8         */
    int32_t flux_splitting;        /* OCFD_SPLIT_SW / OCFD_SPLIT_LLF            */
    int32_t if_scheme_character;   /* IF_Scheme_Character                       */
    int32_t if_viscous;            /* IF_Viscous                                */
    int32_t ibc;                   /* BoundaryCondition code                    */
    int32_t if_mass_force;         /* IF_Mass_Force                             */
    int32_t scheme_boundary[6];    /* Scheme_boundary(1:6)                      */
    int32_t ghost_cell[6];         /* Ghost_Cell(1:6)                           */
    int32_t nfiltering;            /* NFiltering (<= 10)                        */
    int32_t istep_show;            /* Istep_Show                                */
    int32_t precision;             /* OCFD_PREC_*                               */
    int32_t metric_mode;           /* OCFD_METRIC_*                             */
    int32_t reserved_i[8];
    double ma, re, gamma, pr, ref_amu_t0, dt;
    double ud7l_diss;              /* UD7L_Diss                                 */
    double mass_force[3];
    double bc_para[100];           /* BC_Para(1:100)                            */
    double hybrid_para[100];       /* Hybrid_para(1:100)                        */
    double filter_para[10][100];   /* Filter_Para(1:100, m) at [m-1][0..99]     */
    double reserved_d[8];
} ocfd_config;

typedef struct ocfd_handle ocfd_handle;

/* ---- lifetime --------------------------------------------------------- */
/* Replaces allocate_flow_data / allocate_inviscous_data /
 * allocate_vicous_data (OCFD_NS_Solver.f90:93-151, OCFD_inviscous.f90:15-40,
 * OCFD_viscous.f90:11-20): allocates every device array of one block. */
int ocfd_create(const ocfd_config *cfg, int device, ocfd_handle **out);
/* Replaces the deallocate_* routines. */
int ocfd_destroy(ocfd_handle *h);
/* Text of the last error on this thread (never NULL). */
const char *ocfd_last_error(void);
/* Block until all work queued on the handle's stream has finished. */
int ocfd_synchronize(ocfd_handle *h);

/* ---- grid / metrics (init_mesh, OCFD_mesh.f90:1-58) -------------------- */
/* Interior coordinates Axx,Ayy,Azz (nx*ny*nz each, Fortran order) plus the
 * periodic spans Periodic_I/J/KSpan (9 doubles): the device fills the halos,
 * applies updata_Axyz_periodic (OCFD_mesh.f90:109-185), computes
 * comput_Jacobian3d (OCFD_mesh.f90:61-107) with OCFD_dx0/dy0/dz0, the metric
 * halos, Jac_Ghost_boundary (OCFD_GhostCell.f90:1-160, 290-391) and the
 * normalised metrics Akx1=Akx/Ajac ... (OCFD_mesh.f90:41-55). */
int ocfd_set_grid(ocfd_handle *h, const double *x, const double *y, const double *z,
                  const double spans[9]);
/* Upload metrics computed elsewhere: 19 halo'd arrays in the order
 * Akx,Aky,Akz,Aix,Aiy,Aiz,Asx,Asy,Asz,Ajac,Akx1,Aky1,Akz1,Aix1,Aiy1,Aiz1,
 * Asx1,Asy1,Asz1 (the state of flow_data after init_mesh). */
int ocfd_upload_metrics(ocfd_handle *h, const double *const metrics[19]);
/* OCFD_METRIC_UNIFORM only: the ten constant values Akx,Aky,Akz,Aix,Aiy,Aiz,
 * Asx,Asy,Asz,Ajac (the reference's GRID_AND_JACOBIAN3D input holding one
 * value per field); normalised metrics are formed as Akx/Ajac ... */
int ocfd_set_uniform_metrics(ocfd_handle *h, const double m[10]);
int ocfd_download_metrics(ocfd_handle *h, double *const metrics[19]);

/* ---- state ------------------------------------------------------------ */
/* Upload primitives (halo'd d,u,v,w,T; the halo content is copied too) and
 * build f from them as init3d does (OCFD_init.f90:10-20); Istep, tt set the
 * step counter and time. */
int ocfd_upload_primitives(ocfd_handle *h, const double *d, const double *u,
                           const double *v, const double *w, const double *T,
                           int32_t istep, double tt);
/* Upload / download the conserved variables f(nx,ny,nz,5). */
int ocfd_upload_f(ocfd_handle *h, const double *f);
int ocfd_download_f(ocfd_handle *h, double *f);
int ocfd_download_fn(ocfd_handle *h, double *fn);
int ocfd_download_du(ocfd_handle *h, double *du);
int ocfd_upload_du(ocfd_handle *h, const double *du);
/* Snapshot / restart files in the reference's format: save_flow_data
 * (OCFD_IO.f90:44-64) writes record (Istep, tt) then, for d, u, v, w, T, one
 * record per k plane of nx*ny doubles; path NULL names the file
 * OCFDnnnnnnnn.dat after the handle's Istep.  with_f != 0 appends f(:,:,k,m)
 * records (the reference's reader stops before them), so that a restart
 * continues bit for bit.  read_flow_data (OCFD_IO.f90:5-39; path NULL =
 * "opencfd.dat") replaces the interior of the primitives, sets Istep and tt,
 * and takes f from the file when present (*has_f = 1), else recomputes it
 * from the primitives as at start-up; call ocfd_bc afterwards, as the
 * reference's start-up does (OCFD_NS_Solver.f90:15). */
int ocfd_save_flow_data(ocfd_handle *h, const char *path, int32_t with_f);
int ocfd_read_flow_data(ocfd_handle *h, const char *path, int32_t *has_f);
/* Download halo'd primitives (any pointer may be NULL). */
int ocfd_download_primitives(ocfd_handle *h, double *d, double *u, double *v,
                             double *w, double *T);
int ocfd_download_rhybrid(ocfd_handle *h, double *rhybrid);
int ocfd_get_time(ocfd_handle *h, int32_t *istep, double *tt);

/* ---- one reference routine each --------------------------------------- */
/* exchange_boundary_xyz(d,u,v,w,T) (OCFD_NS_Solver.f90:35-39,
 * OCFD_mpi.f90:183-233): on one block the periodic halo copy. */
int ocfd_exchange_primitives(ocfd_handle *h);
/* comput_Rhybrid (OCFD_SetHybrid.f90:17-152).  counts (may be NULL) gets the
 * number of faces using scheme 1/2/3 and the total, as printed at
 * OCFD_SetHybrid.f90:115-135. */
int ocfd_comput_rhybrid(ocfd_handle *h, int64_t counts[4]);
/* du_inviscous (OCFD_inviscous.f90:53-203) */
int ocfd_du_inviscous(ocfd_handle *h);
/* du_inviscous_Character (OCFD_inviscous_character.f90:3-159) */
int ocfd_du_inviscous_character(ocfd_handle *h);
/* du_viscous (OCFD_viscous.f90:29-131) */
int ocfd_du_viscous(ocfd_handle *h);
/* OCFD_time_adv_RK3(KRK) (OCFD_Time_Adv.f90:4-36). */
int ocfd_time_adv_rk3(ocfd_handle *h, int32_t krk);
/* fn = f (OCFD_NS_Solver.f90:23-31) */
int ocfd_save_fn(ocfd_handle *h);
/* comput_duvwT (OCFD_Time_Adv.f90:40-63): returns the number of cells with
 * T <= 0 (> 0 is a numerical event; those cells are repaired as
 * handle_NegativeT does, OCFD_handle_NegativeT.f90:1-17). */
int ocfd_comput_duvwt(ocfd_handle *h);
/* OCFD_bc (OCFD_BC.f90:2-99) incl. Flow_Ghost_boundary. */
int ocfd_bc(ocfd_handle *h);
/* Stage fusion (no counterpart in the reference, whose routines each sweep
 * the whole block).  For a caller that issues the routines above in
 * NS_solver's order each stage (OCFD_NS_Solver.f90:33-60: exchange,
 * comput_Rhybrid, du_inviscous[_Character], du_viscous, OCFD_time_adv_RK3,
 * comput_duvwT, OCFD_bc), on = 1 merges them into ocfd_step's stage
 * kernels, with the same results bit for bit:
 *   ocfd_du_viscous      forms only the viscous fluxes;
 *   ocfd_time_adv_rk3    adds their divergence, applies RK3, stores fn = f
 *                        at KRK = 1, recovers the primitives, refreshes f on
 *                        the boundary planes of a periodic block (ibc 0) and
 *                        lists the T <= 0 cells, in one kernel;
 *   ocfd_save_fn         does nothing (ocfd_time_adv_rk3(1) stores fn);
 *   ocfd_comput_duvwt    repairs the listed cells and returns their count;
 *   ocfd_bc              on a periodic block, only the ghost cells.
 * Arrays read between the calls of one stage differ from the per-routine
 * path's: after ocfd_du_viscous du holds no viscous terms, and after
 * ocfd_time_adv_rk3 f is already refreshed on the boundary planes.  on = 0
 * (the default) restores one kernel sequence per routine. */
int ocfd_set_stage_fusion(ocfd_handle *h, int32_t on);
/* Data of BoundaryCondition = BC_BoundaryLayer (ibc 2, ocfd_bc_BoundaryLayer,
 * OCFD_BC_Boundarylayers.f90:17-167) or BC_SweptCorner (ibc 3,
 * ocfd_bc_swept_corner, :443-617), which the reference prepares on the host
 * in init_bcdata_boundarylayer / init_bc_nonRef / init_bc_wall_perturbation
 * (BC%flow_inlet*, BC%flow_upper*, BC%NonReflect_upper, BC%wall_pertb):
 *   aoa     Para%AoA (radians): free stream (cos aoa, sin aoa, 0)
 *   inlet   d,u,v,w,T at i = 1 as [5][nz][ny] (1-d data repeated over k);
 *           NULL for the free-stream inlet (BC_para(1) = 0)
 *   upper   d,u,v,w,T at j = ny as [5][nz][nx]; NULL for the free stream
 *   nonref  [nz][nx], 1 = non-reflecting upper point (no Dirichlet value);
 *           NULL = all Dirichlet
 *   pertb   [nz][nx] wall blow-and-suction shape; NULL = none
 *   wall    swept corner only, [nz][nx]: 1 = no-slip wall point, 0 = point
 *           extrapolated from j = 2 (OCFD_BC_Boundarylayers.f90:517); NULL =
 *           all wall
 * The other parameters come from the config's bc_para.  Ghost_Cell = -1 on
 * j- (boundary layer) / k- (swept corner) selects the user ghost layers of
 * OCFD_GhostCell_UserDef.f90.  Call before the first ocfd_bc. */
int ocfd_set_bc_boundarylayer(ocfd_handle *h, double aoa, const double *inlet, const double *upper,
                              const int32_t *nonref, const double *pertb, const int32_t *wall);
/* Time factor ht of the wall perturbation for the coming stages
 * (get_ht_multifrequancy, OCFD_BC_Boundarylayers.f90:401-440, evaluated by the
 * caller at the step's time; 1 until set).  Steps replayed from a graph read
 * the current value. */
int ocfd_set_bc_time_factor(ocfd_handle *h, double ht);
/* filtering(f) (OCFD_filtering.f90:4-70) at the handle's current Istep. */
int ocfd_filtering(ocfd_handle *h);
/* dt = cfl / max over the cells of sum over the axes of
 * (|u . grad xi| + c |grad xi|) / h_xi, c = sqrt(T)/Ma, from the device state
 * and metrics (a wavefront / workgroup max-reduction).  No counterpart in the
 * reference, which advances with the fixed Para%dt (OCFD_Time_Adv.f90:20, 30);
 * a caller computes dt once with it and puts it in ocfd_config.dt.  The
 * handle's dt is not changed. */
int ocfd_cfl_dt(ocfd_handle *h, double cfl, double *dt);
/* Advance Istep by one and tt by dt (OCFD_NS_Solver.f90:64-65). */
int ocfd_advance_clock(ocfd_handle *h);

/* ---- whole steps ------------------------------------------------------ */
/* nsteps iterations of the time loop body (OCFD_NS_Solver.f90:23-67): fn=f,
 * three RK stages, Istep/tt update and filtering.  State stays on device. */
int ocfd_step(ocfd_handle *h, int32_t nsteps);

/* ---- sweep tiles (no counterpart in the reference) --------------------- */
/* Tile shape of the inviscid sweep of each axis: 0 = x 64x4, y/z 32x8 points
 * x lines (the default); 1 = x 32x8, y/z 16x16; 2 = x 128x2, y/z 64x4;
 * 3 and 4 (y and z only) march along the axis, one lane per line, over
 * segments of 32 and 64 points, keeping the split fluxes in registers.
 * Shapes 1-4 apply to periodic axes with OCFD_METRIC_UNIFORM whose extents
 * fit them; other axes use 0.  Results are bit-identical for every shape. */
int ocfd_set_sweep_tiles(ocfd_handle *h, const int32_t shape[3]);
int ocfd_get_sweep_tiles(ocfd_handle *h, int32_t shape[3]);
/* Autotuner: times `reps` launches of each axis' sweep in every shape it
 * supports on the block's current state and keeps the fastest (shape_out may
 * be NULL).  Overwrites du (a work array): call it between time steps. */
int ocfd_autotune(ocfd_handle *h, int32_t reps, int32_t shape_out[3]);

/* ---- diagnostics ------------------------------------------------------ */
/* show_flow_msg sums (OCFD_Time_Adv.f90:65-90): out[0..2] = E, K, S
 * (averaged over the global cell count). */
int ocfd_show_flow_msg(ocfd_handle *h, double out[3]);
/* ocfd_show_flow_msg that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The partial sums live in the library's file-scope device array of 16 MiB
 * (2^21 doubles, held by every GPU the library runs on until the process
 * ends), so blocks of at most about 178.9 million cells (3 doubles per 256
 * cells) are accepted; larger ones fail with a message.  Results equal
 * ocfd_show_flow_msg's bit for bit.  All calls on one GPU share the array:
 * calls that overlap on one GPU race. */
int ocfd_show_flow_msg_noalloc(ocfd_handle *h, double out[3]);
/* ana_box statistics (OCFD_ANA.f90:178-311): out[0..19] in the column order
 * of statistics.dat: tt, Re_lamda, Re_lamdax, Amt, vv_rms, u_rms, v_rms,
 * w_rms, u_skewness, u_flatness, ux_skewness, ux_flatness, E_kinetic,
 * Enstrophy, d_rms, T_aver, Amu_aver, d_aver, Ux_rms, Alamdax. */
int ocfd_ana_box(ocfd_handle *h, double out[20]);

/* ---- batches of independent boxes ------------------------------------- */
/* A batch advances many independent blocks ("boxes") of uneven size in the
 * same time steps: each kernel phase of a step (halo fill, the three sweeps
 * per scheme group, viscous fluxes, stage update, T <= 0 repair) is one
 * launch over the tiles of all boxes, and whole steps replay from a captured
 * hipGraph.  Boxes must be periodic on every axis with OCFD_METRIC_UNIFORM,
 * without hybrid scheme, boundary hook, ghost cells or filter; each box keeps
 * its own scheme, splitting, viscous order, Ma, Re, dt; all boxes share one
 * precision.  Results equal stepping each box alone bit for bit.  No
 * counterpart in the reference (one reference run advances one block). */
typedef struct ocfd_batch ocfd_batch;
int ocfd_batch_create(const ocfd_config *cfgs, int32_t nbox, int32_t device, ocfd_batch **out);
int ocfd_batch_destroy(ocfd_batch *b);
/* The handle of box i (owned by the batch, do not destroy): set-up
 * (ocfd_set_uniform_metrics, ocfd_upload_primitives, ocfd_bc), downloads and
 * diagnostics of the box go through the single-block entry points. */
int ocfd_batch_box(ocfd_batch *b, int32_t i, ocfd_handle **h);
/* nsteps time steps of every box; returns the number of T <= 0 cells
 * repaired (error if a box has more than 10 in one stage). */
int ocfd_batch_step(ocfd_batch *b, int32_t nsteps);
/* Autotuner: per scheme group and axis, the fastest sweep tile shape of the
 * batched launch, and the segment length (4, 8, 16 or 32 planes per
 * workgroup) of the k-marching viscous kernels (overwrites the boxes' du and
 * Ev work arrays; results do not depend on either choice). */
int ocfd_batch_autotune(ocfd_batch *b, int32_t reps);
int ocfd_batch_set_sweep_tiles(ocfd_batch *b, int32_t group, const int32_t shape[3]);
/* Number of scheme groups, their tile shapes (3 per group), the total cell
 * count and the marching segment length (any pointer may be NULL). */
int ocfd_batch_info(ocfd_batch *b, int32_t *ngroups, int32_t *shapes, int64_t *cells, int32_t *kseg);
/* How ocfd_batch_step launches a step.  mode 0 (the default): one launch
 * per kernel phase over the tiles of all boxes.  mode 1: one persistent
 * launch per time step: a fixed grid of workgroups takes the (box, phase,
 * tile) items of the step's three stages from a device work queue in phase
 * order, each item waiting only for its own box's previous phase (per-box
 * completion counters), so boxes overlap their phases and no barrier over
 * all boxes remains.  Mode 1 runs x tiles of 64 x 4 and y / z marches of
 * 32-point segments (not the autotuned shapes) and needs non-characteristic
 * schemes and CD6 in every box.  Mode 1 is currently switched off (the call
 * fails): its kernel did not finish on the GPU. */
int ocfd_batch_set_launch_mode(ocfd_batch *b, int32_t mode);

/* ---- instrumentation -------------------------------------------------- */
/* Enable per-kernel HIP-event timing on the handle's stream. */
int ocfd_profile_enable(ocfd_handle *h, int32_t on);
/* Per-kernel accumulated time (ms) and launch count; names are the kernel
 * families ("inviscid_x", ...).  Returns the number of entries written. */
int ocfd_profile_read(ocfd_handle *h, int32_t max_entries, char names[][32],
                      double *ms, int64_t *launches);
int ocfd_profile_reset(ocfd_handle *h);

#ifdef __cplusplus
}
#endif
#endif /* OCFD_HIP_H */
