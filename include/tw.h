/*
 * tw.h -- C-ABI of the MI355X-native staggered SBP-SAT wave-solver hot path.
 *
 * The reference (wsmoses/tpu-wave) exposes no library: its callable surface is
 * the C++ member functions that its three main() drivers call.  Each entry point
 * below replaces one of those, with the reference function it stands in for
 * cited beside it (paths relative to the reference tree).
 *
 * Conventions
 *   - Data axis x is the slow (row) axis, y the fast (column) axis; physical
 *     'Y' (depth) maps to data x and physical 'X' to data y, as in
 *     cuda/src/cpp/namespace_forward.hpp:56-61 (XY_to_01).
 *   - Staggered grid types (type along x, type along y): 'N' = nodes (M+1
 *     points), 'M' = cell centres (M points).  Grid ids follow the reference's
 *     Array_Grid_types order NN, MN, NM, MM (namespace_forward.hpp:159-167).
 *   - Fields are named by data axes:  TW_VX lives on NM (the reference's "Vy"
 *     grid), TW_VY on MN (the reference's "Vx" grid), TW_SXX/TW_SYY on MM
 *     (SMM soln[0]/soln[1]), TW_SXY on NN, TW_P on MM (acoustic).
 *   - Every function returns 0 on success and a nonzero TW_E* code otherwise;
 *     nothing calls exit().  tw_last_error() gives the message.
 *   - A solver handle owns all of its device memory; caller arrays are
 *     borrowed for the duration of the call only.  A handle is not
 *     thread-safe; distinct handles are independent.
 *   - Host arrays are dense row-major [rows][cols] doubles of the field's
 *     grid size (see tw_field_shape).
 */
#ifndef TW_H
#define TW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tw_solver tw_solver;

/* ---- error codes ---- */
#define TW_OK          0
#define TW_EINVAL      1   /* bad argument / inconsistent configuration          */
#define TW_ERANGE      2   /* index or location outside the grid / weak boundary  */
#define TW_ENOMEM      3   /* allocation failed or over the size guard           */
#define TW_EDEVICE     4   /* HIP runtime error                                  */
#define TW_ENONFINITE  5   /* non-finite value detected in the state             */
#define TW_ESTATE      6   /* call not valid in the current state                */

/* ---- enums ---- */
enum tw_physics   { TW_ELASTIC = 0, TW_ACOUSTIC = 1 };
enum tw_precision { TW_FP64 = 0, TW_FP32 = 1, TW_BF16 = 2, TW_FP16 = 3 };
/* compensation variants of the leapfrog update S <- S + R
 * (cuda/src/cuda/grid_update.cut:62-124, cuda/src/cpp/grid_update.tpp:1-128):
 *   TW_CPST_0   S += R, R = 0                      ('0')
 *   TW_CPST_3R  fast2sum(S,R) -> (S,R)             ( 3,'R')
 *   TW_CPST_M3R explicit 3-flop form of the above  (-3,'R')
 *   TW_CPST_6R  2Sum(S,R) -> (S,R)                 ( 6,'R')
 *   TW_CPST_3C  fast2sum(S,R+C) -> (S,C), R = 0    ( 3,'C')
 *   TW_CPST_M3C explicit 3-flop form of the above  (-3,'C')
 *   TW_CPST_6C  2Sum(S,R+C) -> (S,C), R = 0        ( 6,'C')
 *   TW_CPST_K   16-bit storage only: FP32 sum, the rounding residual of the
 *               store is kept in a 16-bit companion array (Kahan-style)
 *   TW_CPST_K8  as K with the residual kept at 8 bits: a code q in
 *               [-127, 127] per value, residual = q * ulp(stored value) / 256
 *   TW_CPST_K8S as K8, and every stencil, SAT and boundary-value input reads
 *               the stored value plus its decoded residual (the compensated
 *               value) instead of the stored value alone */
enum tw_cpst {
  TW_CPST_0 = 0, TW_CPST_3R = 1, TW_CPST_M3R = 2, TW_CPST_6R = 3,
  TW_CPST_3C = 4, TW_CPST_M3C = 5, TW_CPST_6C = 6, TW_CPST_K = 7, TW_CPST_K8 = 8,
  TW_CPST_K8S = 9
};
enum tw_field { TW_VX = 0, TW_VY = 1, TW_SXX = 2, TW_SYY = 3, TW_SXY = 4, TW_P = 5 };
enum tw_grid  { TW_GRID_NN = 0, TW_GRID_MN = 1, TW_GRID_NM = 2, TW_GRID_MM = 3 };
/* material coefficients as used by the update (cuda/src/cpp/grid_parameter.cpp:75-162) */
enum tw_coef  {
  TW_C_BINV_VX = 0,  /* 1/rho on NM            */
  TW_C_BINV_VY = 1,  /* 1/rho on MN            */
  TW_C_MU      = 2,  /* mu on NN (elastic)     */
  TW_C_LAMBDA  = 3,  /* lambda on MM (elastic) */
  TW_C_TWOMU   = 4,  /* 2 mu on MM (elastic)   */
  TW_C_KAPPA   = 5   /* kappa = rho vp^2 on MM (acoustic) */
};

typedef struct tw_config {
  int32_t physics;        /* TW_ELASTIC or TW_ACOUSTIC                               */
  int32_t mx, my;         /* cells along data x (rows) and data y (columns)          */
  double  dx;             /* grid spacing (both axes)                                */
  int64_t num_dx, den_dx; /* dx as num/den for rational locations; den_dx <= 0: none */
  double  dt;             /* time step                                               */
  char    bc[4];          /* boundary of sides xL, xR, yL, yR: 'P' periodic (both
                             sides of an axis), 'F' free-surface SAT,
                             'C' characteristic (absorbing) SAT                      */
  int32_t precision;      /* field storage type (tw_precision); arithmetic is FP64
                             for TW_FP64 and FP32 otherwise                          */
  int32_t coef16;         /* BF16/FP16 only: material coefficient storage.
                             0: FP32; 1: the 16-bit storage type; 2: one byte
                             per value, each row mapped affinely onto codes
                             0..255 (FP32 offset = row minimum, step = row
                             range / 255, value = offset + code * step in FP32);
                             3: the same with two-byte codes 0..65535 (step =
                             row range / 65535).
                             tw_get_coef returns the values the kernels use.   */
  int32_t cpst;           /* tw_cpst                                                 */
  int32_t stencil_form;   /* 0: stencil*(dt/dx) (cuda/src/cpp/grid.hpp:281-282);
                             1: stencil*dt/dx (singlecore/src/cpp/namespaces.hpp:607-608) */
  int32_t energy_form;    /* 0: H-norm energy with compliance weights, /2
                             (cuda/src/cpp/grid_energy.cpp:8-216);
                             1: acoustic reference sum soln*prmt*soln*w with
                             w = dx^2/2 (halved at node ends), no /2
                             (singlecore/src/cpp/namespaces.hpp:734-777,
                              singlecore/src/main_cpp/cpu_acoustic.cpp:327-344)      */
  int32_t record_energy;  /* 1: record energy every step                             */
  int32_t nt_max;         /* capacity (steps) of the receiver and energy records     */
  double  source_scaling; /* source increment is divided by it
                             (cuda/src/cuda_main/nvcc_device_FDM_2D.cu:554)          */
  int32_t device;         /* HIP device ordinal                                      */
} tw_config;

/* ---- host-only helpers (no device needed) ---- */

/* Fills cfg with defaults: elastic, free-surface x / periodic y, FP32, no
 * compensation, energy on.  Sizes, dx, dt and nt_max must then be set. */
void tw_config_default(tw_config* cfg);

/* Checks a configuration; 0 if usable. Message via tw_global_error(). */
int tw_config_check(const tw_config* cfg);

/* Shape of a field's grid: rows = mx or mx+1, cols = my or my+1 by type. */
int tw_field_shape(const tw_config* cfg, int32_t field, int32_t* rows, int32_t* cols);

/* Rational physical location -> grid index on a grid type ("MN" etc. as
 * data-x type, data-y type).  Replaces the index maps of
 * Class_Forward_Specs::process_src_locations / process_rcv_locations
 * (cuda/src/cpp/forward.cpp:30-54, 127-151): node index num*den_dx/(den*num_dx),
 * cell index (2*num*den_dx - den*num_dx)/(2*den*num_dx); the division must be
 * exact, otherwise TW_ERANGE. */
int tw_locate(const tw_config* cfg, int32_t grid, int64_t num_x, int64_t den_x,
              int64_t num_y, int64_t den_y, int32_t* ix, int32_t* iy);

/* Bilinear interpolation of a node-based model (nxm+1) x (nym+1) with spacing
 * num_dxm/den_dxm onto the points of one staggered grid.  Replaces
 * Class_Grid::interpolate_forward_parameter (cuda/src/cpp/grid_parameter.cpp:6-60). */
int tw_interpolate_model(const tw_config* cfg, int32_t grid, const double* model,
                         int32_t nxm, int32_t nym, int64_t num_dxm, int64_t den_dxm,
                         double* out);

/* Transpose of tw_interpolate_model: model_out[(nxm+1)][(nym+1)] = the sum
 * over the grid's points of each bilinear weight times values[point], so that
 * gradients on a staggered grid become gradients on the model nodes.  No
 * counterpart in the reference (it has no gradient); it transposes
 * Class_Grid::interpolate_forward_parameter (cuda/src/cpp/grid_parameter.cpp:6-60)
 * for the model parameters of Class_Inverse_Specs (cuda/src/cpp/inverse.hpp:10). */
int tw_interpolate_model_adjoint(const tw_config* cfg, int32_t grid, const double* values,
                                 int32_t nxm, int32_t nym, int64_t num_dxm, int64_t den_dxm,
                                 double* model_out);

/* Ricker wavelet increment A*dt/dx^2/source_scaling for a step; replaces the
 * source evaluation in forward_simulation_periodic_y
 * (cuda/src/cpp/forward_simulation.cpp:62-70,104-112). half = 0 for velocity
 * grids (t = it*dt - t0), 1 for stress/pressure grids (t = (it+1/2)*dt - t0). */
double tw_ricker_increment(const tw_config* cfg, double f0, double t0, int32_t it, int32_t half);

const char* tw_global_error(void);
const char* tw_version(void);

/* ---- solver lifecycle ---- */

/* Grid setup: replaces Class_Grid::set_grid_parameters / set_forward_operators /
 * set_grid_pointers (cuda/src/cpp/grid.hpp:146-481) and
 * cuda_Class_Grid::cuda_Class_Grid_initialize (cuda/src/cuda/grid.cuh:216-408).
 * Fields start at zero; coefficients default to rho = kappa = 1, vp = 2, vs = 1
 * (the reference's verification material, nvcc_device_FDM_2D.cu:153-159). */
int tw_create(const tw_config* cfg, tw_solver** out);
int tw_destroy(tw_solver* s);
const char* tw_last_error(const tw_solver* s);

/* Material from a node model: replaces retrieve_forward_parameter,
 * define_parameters_energy, adjust_parameters_energy_periodic and
 * make_density_reciprocal (cuda/src/cpp/grid_parameter.cpp:75-162,
 * cuda/src/cpp/grid_energy.cpp:8-177).  Arrays are (nxm+1) x (nym+1) nodes;
 * for TW_ACOUSTIC vs is ignored and kappa = rho*vp^2. */
int tw_set_model(tw_solver* s, const double* rho, const double* vp, const double* vs,
                 int32_t nxm, int32_t nym, int64_t num_dxm, int64_t den_dxm);

/* One coefficient array on its grid (tw_coef), overriding tw_set_model's. */
int tw_set_coef(tw_solver* s, int32_t coef, const double* values);
int tw_get_coef(tw_solver* s, int32_t coef, double* values);

/* Point source on every field of a grid (cuda/src/cpp/forward.cpp:10-100 for
 * the placement rules; forward_simulation.cpp:57-114 for the injection).
 * TW_ERANGE if off-grid, or on an x node boundary of an SBP axis. */
int tw_set_source_ricker(tw_solver* s, int32_t grid, int32_t ix, int32_t iy,
                         double f0, double t0);
/* Custom source: increment per step (already scaled), n >= steps run. */
int tw_set_source_wavelet(tw_solver* s, int32_t grid, int32_t ix, int32_t iy,
                          const double* increments, int32_t n);
int tw_clear_source(tw_solver* s);

/* Receivers, one field each (Map_grid_struct_rcv_forward,
 * cuda/src/cpp/forward.cpp:104-211). Replaces any previous set. */
int tw_set_receivers(tw_solver* s, int32_t n, const int32_t* field,
                     const int32_t* ix, const int32_t* iy);
/* The registered receivers in record order: *n = their count; with non-null
 * arrays (capacity cap) also their fields and indices (TW_ERANGE if cap < n). */
int tw_get_receivers(tw_solver* s, int32_t cap, int32_t* n, int32_t* field, int32_t* ix, int32_t* iy);

/* Advance nsteps leapfrog steps: one iteration of the time loop of
 * forward_simulation_periodic_y (cuda/src/cpp/forward_simulation.cpp:30-179) /
 * nvcc_device_FDM_2D.cu:449-730 per step.  Asynchronous w.r.t. the host. */
int tw_step(tw_solver* s, int32_t nsteps);
int tw_synchronize(tw_solver* s);
int tw_current_step(tw_solver* s, int32_t* it);

/* Receiver traces [n][nt] (Map_grid_RESULT_rcv, nvcc_device_FDM_2D.cu:737-751)
 * and energy records (Record_E_k/E_p0/E_p1, forward.cpp:288-327), nt = steps
 * run since the last reset; multiplied back by source_scaling as the
 * reference's writers do. Null pointers are skipped. */
int tw_read_receivers(tw_solver* s, double* out, int32_t* nt);
int tw_read_energy(tw_solver* s, double* ek, double* ep0, double* ep1, int32_t* nt);

/* Field snapshot / initial condition (cuda_Class_Grid::copy_field_pitched,
 * cuda/src/cuda/grid.cuh:414-512).  Values are in stored units (not rescaled). */
int tw_get_field(tw_solver* s, int32_t field, double* out);
int tw_set_field(tw_solver* s, int32_t field, const double* in);
/* Companion array (R or C compensation, or the Kahan residual) of a field;
 * for TW_CPST_K8 the residual codes as integers (-127..127). */
int tw_get_companion(tw_solver* s, int32_t field, double* out);
/* The field's value including what its compensation carries: the FP64 sum of
 * the stored value and its R, C or K companion, or of the stored value and
 * its decoded K8 residual (q * ulp(value) / 256); the stored value for
 * TW_CPST_0.  This is the value the compensated update works with (the
 * reference keeps S and its error term side by side,
 * cuda/src/cuda/grid_update.cut:62-124), and the one to compare with a
 * higher-precision run. */
int tw_get_field_compensated(tw_solver* s, int32_t field, double* out);

/* Zero all fields, companions, records and the step counter. */
int tw_reset(tw_solver* s);

/* Checkpoint of the complete dynamic state (fields, companions, step index);
 * restoring it and stepping reproduces an uninterrupted run bit for bit. */
int tw_checkpoint_size(tw_solver* s, size_t* bytes);
int tw_checkpoint_save(tw_solver* s, void* buf, size_t bytes);
int tw_checkpoint_load(tw_solver* s, const void* buf, size_t bytes);

/* Run on the caller's HIP stream (hipStream_t) instead of the handle's own. */
int tw_set_stream(tw_solver* s, void* hip_stream);

/* Check the state for non-finite values (TW_ENONFINITE if any). */
int tw_check_finite(tw_solver* s);
/* tw_check_finite without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The flag lives in a
 * fixed 256-byte device array per device (its size does not depend on the
 * grid, so no input exceeds it); it is cleared on every call, and calls that
 * overlap on one device race for it. */
int tw_check_finite_noalloc(tw_solver* s);

/* Configuration the solver was created with. */
int tw_get_config(const tw_solver* s, tw_config* cfg);

/* ---- reference-format front-end (host only) ----
 * The reference reads its problem from text files in a key_value_value format
 * (cuda/src/cpp/input_processing.cpp:11-137), lets command-line options
 * override them (:140-200), reads sources and receivers from SrcRcv.txt
 * (cuda/src/cpp/inverse.cpp:133-348) and the material from raw FP64 files
 * (inverse.cpp:6-130), and writes one Signal_*.txt per receiver and one
 * Energy_*.txt per source (cuda/src/cpp/forward.cpp:215-327).  These entry
 * points read and write the same formats.  Axis convention as above: the
 * files' physical X is data axis 1 (y), physical Y is data axis 0 (x). */
typedef struct tw_input {
  int32_t model_m[2];              /* model cells along data x, y   (model_<X>_<Y>) */
  int32_t grid_m[2];               /* simulation cells along data x, y (grids_<X>_<Y>) */
  int64_t num_dx_model, den_dx_model; /* model spacing num/den (dxM_<num>_<den>)    */
  int64_t num_dx, den_dx;          /* simulation spacing num/den (dxS_<num>_<den>)  */
  char    bc[4];                   /* xL, xR, yL, yR as tw_config.bc (bdry_<XL><XR><YL><YR>,
                                      physical X first)                              */
  int32_t nt;                      /* Nt_                                            */
  double  dt_max;                  /* MAXdt_: dt = dt_max * cfl                      */
  double  cfl;                     /* CFL_                                           */
  double  central_f;               /* frequency_ (Ricker)                            */
  double  time_delay;              /* delay_ (Ricker t0)                             */
  int32_t device;                  /* device_                                        */
  char    energy;                  /* 'Y' or 'N' (option -energy)                    */
  char    medium[64];              /* option -medium: the data folder name           */
} tw_input;

/* Reference defaults (cuda/src/cpp/namespace_input.hpp:13-68): all sides 'F',
 * Nt -1, dt_max 2^30, CFL 0.625, f 5, t0 0.25, energy 'Y', device -1,
 * medium "homogeneous"; sizes and spacings unset (-1). */
void tw_input_default(tw_input* in);
/* Parse an InputFile.txt (file_input_processing): lines whose first
 * '_'-separated token is a known key set the value; other lines are ignored.
 * TW_EINVAL if the file cannot be opened or a value does not parse. */
int tw_read_input_file(const char* path, tw_input* in);
/* Apply one command-line option as command_line_input_processing does:
 * name is one of medium, B_type, Nt, dt_max, CFL, energy, f, t0, device. */
int tw_input_option(tw_input* in, const char* name, const char* value);
/* checking_and_printouts_input_parameters + ns_input_derived_variables
 * (input_processing.cpp:203-214, namespace_input.hpp:124-146) and the
 * numerical parameters of main_host_FDM_2D.tpp:60-61: checks that model and
 * simulation sizes describe the same extent and that energy is 'Y' or 'N',
 * then fills cfg (defaults of tw_config_default for everything the file does
 * not set) with sizes, dx = num/den, dt = dt_max * cfl, boundaries, nt_max = Nt,
 * device, and the given physics, precision and compensation. */
int tw_input_to_config(const tw_input* in, int32_t physics, int32_t precision, int32_t cpst, tw_config* cfg);
/* Note: the reference's elastic solver runs periodic along data y and with
 * free-surface SAT along data x whatever the letters say (it requires them
 * all to be 'F', namespace_input.hpp:158-160, and differentiates y
 * periodically, grid_drvt.tpp:1-3); the letters are taken here as written, so a
 * caller reproducing a reference run sets cfg->bc to "FFPP" afterwards. */

typedef struct tw_srcrcv_entry {
  int32_t kind;       /* 0 source (Src_ line), 1 receiver (Rcv_ line)            */
  int32_t src_index;  /* source index (a receiver's source)                     */
  int32_t rcv_index;  /* receiver index within its source; -1 for sources        */
  char    grid[2];    /* grid type along data x, y ('N' or 'M')                  */
  int64_t num[2];     /* location along data x, y as num/den in length units     */
  int64_t den[2];
} tw_srcrcv_entry;

/* Parse a SrcRcv.txt (Class_Inverse_Specs::input_SrcRcv_locations): entries in
 * file order.  With out == NULL only *n is set.  Errors as the reference's
 * checks: a receiver before its source, duplicate source locations or indices,
 * duplicate receiver locations or indices within a source (TW_EINVAL); more
 * entries than cap (TW_ERANGE, *n = the count needed). */
int tw_read_srcrcv_file(const char* path, tw_srcrcv_entry* out, int32_t cap, int32_t* n);

/* Read a raw FP64 model file (rho, vp or vs) of (mx_model+1) x (my_model+1)
 * nodes, data x outer (Class_Inverse_Specs::input_inverse_parameter,
 * inverse.cpp:6-130): the byte size must match exactly; on an axis whose two
 * sides are 'P' in bc the first and last node lines must be equal. */
int tw_read_model_file(const char* path, int32_t mx_model, int32_t my_model, const char bc[4], double* out);

/* Set up one shot of a SrcRcv list on a solver (process_src_locations and
 * process_rcv_locations, forward.cpp:10-211): a Ricker source (central_f,
 * time_delay of in) at the source's location, and every field of each of its
 * receivers' grids recorded, receivers in file order, fields in the
 * reference's order per grid (NN: sxy; MN: vy; NM: vx; MM: sxx then syy, or p).
 * The checks of tw_set_source_ricker and tw_set_receivers apply (no source on
 * a weak node boundary, forward.cpp:95-98); a source index that is not in the
 * list or has no receivers is refused (TW_EINVAL, forward.cpp main loop
 * main_host_FDM_2D.tpp:203-205). */
int tw_setup_shot(tw_solver* s, const tw_input* in, const tw_srcrcv_entry* entries, int32_t n, int32_t src_index);
/* Write the results of the shot set up with tw_setup_shot after tw_step:
 * Signal_src_<i>_<grid>_<iy>_<ix>_rcv_<j>_<grid>_<iy>_<ix>.txt per receiver
 * (grid letters physical X first; each line "% 6d  " then "% 16.15e  " per
 * field, scaled by source_scaling) and, when energy is recorded,
 * Energy_src_<i>_<grid>_<iy>_<ix>.txt with E = E_k + (E_p0 + E_p1)/2 scaled by
 * source_scaling^2, "%16.15e" per line.  folder must exist.  The solver
 * must record exactly this shot's receivers (as tw_setup_shot registers
 * them), otherwise TW_ESTATE. */
int tw_write_results(tw_solver* s, const char* folder, const tw_srcrcv_entry* entries, int32_t n, int32_t src_index);
/* Final-state snapshots as the reference's device driver writes them
 * (cuda/src/cuda_main/nvcc_device_FDM_2D.cu:788-819): for each grid in the
 * order NN, MN, NM, MM and each of its fields, Snapshot_<grid type, physical
 * X first>_<Sxy|Vx|Vy|SMM>_<field index on the grid>.txt with the stored
 * values (no compensation added) times source_scaling, "% 16.15e" per line,
 * row-major over the grid (data x outer).  Acoustic: P on the MM grid, vx on
 * NM, vy on MN.  folder must exist. */
int tw_write_snapshots(tw_solver* s, const char* folder);
/* Result folder of the reference's device driver
 * (cuda/src/cuda_main/nvcc_device_FDM_2D.cu:755-772):
 * <root>/<elastic|acoustic>/<medium>/device/grid_size_<num>_<den>/Nt_<Nt>/
 * <fp64|fp32|bf16|fp16>/dt_<dt as %16.15e with the '.' as 'p'>/<cpst>, cpst as
 * 0_0, 3_R, -3_R, 6_R, 3_C, -3_C, 6_C, K, K8 or K8S.  Writes at most cap bytes. */
int tw_result_folder(const tw_input* in, const tw_config* cfg, const char* root, char* out, int32_t cap);

/* Data misfit of the recorded traces against data [n][nt] (the solver's
 * receiver order, nt <= recorded steps): per record sum over steps of
 * (trace - data)^2 into per_record[n] (may be NULL) and their sum into
 * *total.  Replaces Class_Forward_Specs::calculate_data_misfit
 * (cuda/src/cpp/forward.cpp:414-445), whose per-receiver value is the sum
 * over the fields of the receiver's grid: with tw_setup_shot's registration a
 * receiver's fields are consecutive records.  Sums in FP64 in step order. */
int tw_data_misfit(tw_solver* s, const double* data, int32_t nt, double* per_record, double* total);

/* ---- per-launch timing of the core kernels ----
 * No counterpart in the reference, whose device driver has no timers
 * (cuda/src/cuda_main/nvcc_device_FDM_2D.cu:400).  After tw_profile_begin,
 * the next max_steps steps record HIP events on the solver's stream around
 * the core launch of each half step (velocity, stress; the single core launch
 * of the fused step counts as velocity).  tw_profile_end waits for the stream
 * and returns the summed durations (ms) and the number of profiled steps. */
int tw_profile_begin(tw_solver* s, int32_t max_steps);
int tw_profile_end(tw_solver* s, double* ms_velocity, double* ms_stress, int32_t* nsteps);

/* Device-to-device streaming copy of `bytes` (a multiple of 16, 16-byte
 * aligned device pointers) on a HIP stream (NULL: the default stream), with
 * 16-byte non-temporal loads and stores: the attainable copy bandwidth the
 * benchmark reports beside the solver's.  No counterpart in the reference. */
int tw_stream_copy(void* dst, const void* src, size_t bytes, void* hip_stream);

/* ---- discrete adjoint (acoustic and elastic; FP64 or FP32 storage; no compensation) ----
 *
 * Gradient of a receiver misfit L(traces) with respect to the material
 * coefficients, by the transposed time step run backwards with sqrt(N)-step
 * checkpointing and forward recomputation.  Stands in for the reference's
 * gradient path of the inversion driver (cuda/src/cpp/inverse.cpp,
 * cuda/src/cpp/forward.cpp:414-445: misfit sum (RESULT - record)^2).
 *
 * tw_adjoint_forward: advance nsteps from the current state like tw_step,
 * keeping a checkpoint every `interval` steps (<= 0: ceil(sqrt(nsteps))).
 * Receiver traces and energies are recorded as by tw_step. */
int tw_adjoint_forward(tw_solver* s, int32_t nsteps, int32_t interval);

/* tw_adjoint_run: given rcv_adjoint[n][nsteps] = dL/d(trace value) for the
 * receivers of the last tw_adjoint_forward run (trace units as returned by
 * tw_read_receivers divided by source_scaling), write dL/d(coefficient) as
 * dense [rows][cols] doubles on each coefficient's grid: kappa (MM), 1/rho on
 * NM, 1/rho on MN (null pointers skipped).  The fields are left as after the
 * forward run, and so are its receiver and energy records (the steps the
 * sweep recomputes record nothing). */
int tw_adjoint_run(tw_solver* s, const double* rcv_adjoint, double* grad_kappa, double* grad_binv_vx,
                   double* grad_binv_vy);

/* tw_adjoint_run_coefs: the same for either physics, with grads[k] the
 * output for coefficient id k (TW_C_*, six entries, null skipped; non-null
 * only for coefficients the physics uses): acoustic kappa and 1/rho on NM and
 * MN; elastic lambda and 2 mu (MM), mu (NN), 1/rho on NM and MN.  Gradients
 * through the characteristic impedances (sqrt(modulus / b) of the normal and,
 * elastic, the tangential velocity) are included.  Like tw_adjoint_run, it
 * differentiates the reference's misfit path (cuda/src/cpp/forward.cpp:414-445)
 * with respect to density and moduli; the reference has no adjoint. */
int tw_adjoint_run_coefs(tw_solver* s, const double* rcv_adjoint, double* const* grads);

/* ---- batched shots ----
 *
 * No counterpart in the reference, which runs its shots one after another
 * (cuda/src/cpp/main_host_FDM_2D.tpp:21-221 loops over sources and calls the
 * forward simulation for each).  A batch groups independent solvers (same
 * physics, storage, compensation and device; any sizes, materials, sources
 * and boundaries) and advances them together: one launch per half step for
 * the whole batch, the time loop captured in a hipGraph and replayed when the
 * same step range is run again (a change of source, receivers or material
 * on any solver rebuilds it).  Batched steps record each solver's energy and
 * receiver samples as tw_step does (same values; energies summed over the
 * batch's tile height, so equal to tw_step's to rounding when the heights
 * differ).  The solvers stay usable on their own; calls on them are ordered
 * after the batch's work. */
typedef struct tw_batch tw_batch;
int tw_batch_create(tw_solver* const* solvers, int32_t n, tw_batch** out);
int tw_batch_destroy(tw_batch* b);
const char* tw_batch_last_error(const tw_batch* b);
/* Advance every solver of the batch by nsteps. */
int tw_batch_step(tw_batch* b, int32_t nsteps);
/* Device time (ms) of the last tw_batch_step's graph launch; waits for it. */
int tw_batch_elapsed(tw_batch* b, double* ms);
/* Rows per block of the core launches (the tile height). */
int tw_batch_set_rows(tw_batch* b, int32_t rows);
/* Time trial_steps batched steps for each candidate tile height (4..64 rows),
 * keep the fastest, and leave the solvers' states and step counters as they
 * were. */
int tw_batch_autotune(tw_batch* b, int32_t trial_steps, int32_t* best_rows);

#ifdef __cplusplus
}
#endif
#endif /* TW_H */
