/*
 * tw_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY, see tw_oracle.h).
 *
 * Shared part: operator constants, configuration checks, grid shapes,
 * rational location maps, bilinear model interpolation, Ricker increments,
 * impedances for the characteristic SAT, and the dispatch to one of four
 * precision instances (tw_oracle_body.h):
 *   f64  FP64 storage, FP64 arithmetic   (reference PRCSFLAG=64)
 *   f32  FP32 storage, FP32 arithmetic   (reference PRCSFLAG=32)
 *   bf16 BF16 storage, FP32 arithmetic   (not in the reference)
 *   f16  FP16 storage, FP32 arithmetic   (not in the reference; its FP16 mode
 *                                         computes in half precision)
 * Build: oracle/Makefile (clang, -O2 -ffp-contract=off, OpenMP).
 */
#include "tw_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>

/* the 8-bit residual variants: K8, and K8S whose stencils read the residual too */
#define OR_K8(cpst) ((cpst) == TW_CPST_K8 || (cpst) == TW_CPST_K8S)
#endif

/* ---- operator constants: cuda/src/cpp/namespace_forward.hpp:65-156 ---- */
static const double OR_STENCIL[4] = {1. / 24, -9. / 8, 9. / 8, -1. / 24};
/* W_intr.DM_L/DM_R: rows of the derivative onto an N-type output near the
 * boundary, columns = first/last five M-type partner points */
static const double OR_DM_L[4][5] = {{-2., 3., -1., 0., 0.},
                                     {-1., 1., 0., 0., 0.},
                                     {1. / 24., -9. / 8., 9. / 8., -1. / 24., 0.},
                                     {-1. / 71., 6. / 71., -83. / 71., 81. / 71., -3. / 71.}};
static const double OR_DM_R[4][5] = {{3. / 71., -81. / 71., 83. / 71., -6. / 71., 1. / 71.},
                                     {0., 1. / 24., -9. / 8., 9. / 8., -1. / 24.},
                                     {0., 0., 0., -1., 1.},
                                     {0., 0., 1., -3., 2.}};
/* W_intr.DN_L/DN_R: derivative onto an M-type output from N-type data */
static const double OR_DN_L[3][5] = {{-79. / 78., 27. / 26., -1. / 26., 1. / 78., 0.},
                                     {2. / 21., -9. / 7., 9. / 7., -2. / 21., 0.},
                                     {1. / 75., 0., -27. / 25., 83. / 75., -1. / 25.}};
static const double OR_DN_R[3][5] = {{1. / 25., -83. / 75., 27. / 25., 0., -1. / 75.},
                                     {0., 2. / 21., -9. / 7., 9. / 7., -2. / 21.},
                                     {0., -1. / 78., 1. / 26., -27. / 26., 79. / 78.}};
/* A_diag (H-norm diagonal at the closure, unit spacing) and projection */
static const double OR_AN_L[4] = {7. / 18., 9. / 8., 1., 71. / 72.};
static const double OR_AN_R[4] = {71. / 72., 1., 9. / 8., 7. / 18.};
static const double OR_AM_L[3] = {13. / 12., 7. / 8., 25. / 24.};
static const double OR_AM_R[3] = {25. / 24., 7. / 8., 13. / 12.};
static const double OR_PM_L[3] = {15. / 8., -5. / 4., 3. / 8.};
static const double OR_PM_R[3] = {3. / 8., -5. / 4., 15. / 8.};

static char or_err[512];
static void or_seterr(const char* m) { snprintf(or_err, sizeof or_err, "%s", m); }
const char* or_error(void) { return or_err; }

void or_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

static int or_field_grid(int physics, int f) {
  (void)physics;
  switch (f) {
    case TW_VX: return TW_GRID_NM;
    case TW_VY: return TW_GRID_MN;
    case TW_SXY: return TW_GRID_NN;
    default: return TW_GRID_MM;
  }
}
static int or_field_coef(int physics, int f) {
  (void)physics;
  switch (f) {
    case TW_VX: return TW_C_BINV_VX;
    case TW_VY: return TW_C_BINV_VY;
    case TW_SXY: return TW_C_MU;
    case TW_P: return TW_C_KAPPA;
    default: return TW_C_LAMBDA;
  }
}
static int grid_tx(int g) { return g == TW_GRID_NN || g == TW_GRID_NM; }
static int grid_ty(int g) { return g == TW_GRID_NN || g == TW_GRID_MN; }
static int coef_grid(int k) {
  switch (k) {
    case TW_C_BINV_VX: return TW_GRID_NM;
    case TW_C_BINV_VY: return TW_GRID_MN;
    case TW_C_MU: return TW_GRID_NN;
    default: return TW_GRID_MM;
  }
}

/* H-norm factor along one axis for point i of n on a grid of type t
 * (adjust_parameters_energy_periodic, cuda/src/cpp/grid_energy.cpp:58-177):
 * SBP axis: (A*dx rounded to the arithmetic type) at the closure, dx inside;
 * periodic axis: dx, with the duplicated last node weighted 0. */
static double or_hweight_d(const tw_config* c, int axis, int t, int i, int n, const double* hN, const double* hM) {
  int periodic = c->bc[2 * axis] == 'P';
  if (periodic) return (t && i == n - 1) ? 0. : c->dx;
  int nb = t ? 4 : 3;
  const double* h = t ? hN : hM;
  if (i < nb) return h[i];
  if (i >= n - nb) return h[n - 1 - i];
  return c->dx;
}

typedef struct or_common or_common;
typedef struct {
  void* (*create)(or_common*);
  void (*destroy)(void*);
  void (*load_coef)(void*, int);
  void (*load_z)(void*);
  void (*set_field)(void*, int, const double*, int);
  void (*get_field)(void*, int, double*, int);
  void (*step)(void*);
  int (*table)(void*, int, double*);
} or_vtbl;

struct or_common {
  tw_config cfg;
  int nf;
  int fields[5];
  int rows[6], cols[6];
  double* coef[6];
  int coef_rows[6], coef_cols[6];
  double* z[6][4];
  int zlen[6][4];
  int src_grid, src_ix, src_iy, src_n;
  double* src_inc;
  int nrcv;
  int *rf, *rix, *riy;
  double* rec;
  double *ek, *ep0, *ep1;
  int it;
  int ep_valid;
  double ep_prev;
};

struct or_solver {
  or_common cm;
  const or_vtbl* vt;
  void* impl;
};

/* ---- storage conversions ---- */
static inline uint16_t f32_to_bf16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u); /* quiet NaN */
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
static inline float bf16_to_f32(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

/* the H-weight helper in the arithmetic type of an instance */
#define or_hweight(c, axis, t, i, n, hN, hM) or_hweight_inst((c), (axis), (t), (i), (n), (hN), (hM))

/* ---- f64 instance ---- */
#define PFX f64
#define SREAL double
#define CREAL double
#define S2C(x) (x)
#define C2S(x) (x)
#define REFMODE 1
static double or_hweight_inst_f64(const tw_config* c, int axis, int t, int i, int n, const double* hN, const double* hM) {
  return or_hweight_d(c, axis, t, i, n, hN, hM);
}
#undef or_hweight
#define or_hweight(c, axis, t, i, n, hN, hM) or_hweight_inst_f64((c), (axis), (t), (i), (n), (hN), (hM))
#include "tw_oracle_body.h"
#undef PFX
#undef SREAL
#undef CREAL
#undef S2C
#undef C2S
#undef REFMODE
#undef or_hweight

/* float instances share a float H-weight helper (hN/hM held as float) */
static double or_hweight_f(const tw_config* c, int axis, int t, int i, int n, const float* hN, const float* hM) {
  double dN[4], dM[3];
  for (int k = 0; k < 4; k++) dN[k] = hN[k];
  for (int k = 0; k < 3; k++) dM[k] = hM[k];
  return or_hweight_d(c, axis, t, i, n, dN, dM);
}
#define or_hweight(c, axis, t, i, n, hN, hM) or_hweight_f((c), (axis), (t), (i), (n), (hN), (hM))

#define PFX f32
#define SREAL float
#define CREAL float
#define S2C(x) (x)
#define C2S(x) ((float)(x))
#define REFMODE 1
#include "tw_oracle_body.h"
#undef PFX
#undef SREAL
#undef CREAL
#undef S2C
#undef C2S
#undef REFMODE

/* exponent field of a 16-bit stored value and the bias that makes
 * 2^(max(e, 1) - bias) = ulp / 256 (the unit of the 8-bit residual code):
 * BF16 ulp = 2^(e - 127 - 7), FP16 ulp = 2^(e - 15 - 10) */
static inline int bf16_exp(uint16_t h) { return (h >> 7) & 0xff; }
static inline int f16_exp(_Float16 h) {
  uint16_t b;
  memcpy(&b, &h, 2);
  return (b >> 10) & 0x1f;
}

#define PFX bf16
#define SREAL uint16_t
#define CREAL float
#define S2C(x) bf16_to_f32(x)
#define C2S(x) f32_to_bf16((float)(x))
#define S_EXP(x) bf16_exp(x)
#define K8_BIAS 142
#define REFMODE 0
#include "tw_oracle_body.h"
#undef PFX
#undef SREAL
#undef CREAL
#undef S2C
#undef C2S
#undef S_EXP
#undef K8_BIAS
#undef REFMODE

#define PFX f16
#define SREAL _Float16
#define CREAL float
#define S2C(x) ((float)(x))
#define C2S(x) ((_Float16)(float)(x))
#define S_EXP(x) f16_exp(x)
#define K8_BIAS 33
#define REFMODE 0
#include "tw_oracle_body.h"
#undef PFX
#undef SREAL
#undef CREAL
#undef S2C
#undef C2S
#undef S_EXP
#undef K8_BIAS
#undef REFMODE
#undef or_hweight

/* ---- configuration ---- */
static int cfg_check(const tw_config* c) {
  if (!c) { or_seterr("null config"); return TW_EINVAL; }
  if (c->physics != TW_ELASTIC && c->physics != TW_ACOUSTIC) { or_seterr("physics"); return TW_EINVAL; }
  if (c->mx < 8 || c->my < 8) { or_seterr("mx, my must be >= 8"); return TW_EINVAL; }
  if (!(c->dx > 0) || !(c->dt > 0)) { or_seterr("dx, dt must be > 0"); return TW_EINVAL; }
  for (int k = 0; k < 4; k++)
    if (c->bc[k] != 'P' && c->bc[k] != 'F' && c->bc[k] != 'C') { or_seterr("bc must be P, F or C"); return TW_EINVAL; }
  for (int a = 0; a < 2; a++)
    if ((c->bc[2 * a] == 'P') != (c->bc[2 * a + 1] == 'P')) { or_seterr("periodic must pair"); return TW_EINVAL; }
  if (c->precision < TW_FP64 || c->precision > TW_FP16) { or_seterr("precision"); return TW_EINVAL; }
  if (c->cpst < TW_CPST_0 || c->cpst > TW_CPST_K8S) { or_seterr("cpst"); return TW_EINVAL; }
  int ref = c->precision == TW_FP64 || c->precision == TW_FP32;
  if (ref && (c->cpst == TW_CPST_K || OR_K8(c->cpst))) { or_seterr("cpst K/K8/K8S needs 16-bit storage"); return TW_EINVAL; }
  if (!ref && c->cpst != TW_CPST_0 && c->cpst != TW_CPST_K && !OR_K8(c->cpst)) {
    or_seterr("16-bit storage supports cpst 0, K or K8");
    return TW_EINVAL;
  }
  if (c->energy_form == 1 && (c->physics != TW_ACOUSTIC || c->bc[0] != 'P' || c->bc[2] != 'P')) {
    or_seterr("energy_form 1 is the periodic acoustic reference form");
    return TW_EINVAL;
  }
  if (c->nt_max < 0) { or_seterr("nt_max"); return TW_EINVAL; }
  if (!(c->source_scaling > 0)) { or_seterr("source_scaling"); return TW_EINVAL; }
  return 0;
}

int or_field_shape(const tw_config* c, int f, int* rows, int* cols) {
  if (f < 0 || f > 5) return TW_EINVAL;
  if (c->physics == TW_ELASTIC && f == TW_P) return TW_EINVAL;
  if (c->physics == TW_ACOUSTIC && (f == TW_SXX || f == TW_SYY || f == TW_SXY)) return TW_EINVAL;
  int g = or_field_grid(c->physics, f);
  *rows = c->mx + grid_tx(g);
  *cols = c->my + grid_ty(g);
  return 0;
}

/* cuda/src/cpp/forward.cpp:30-54 (sources) and 127-151 (receivers) */
static int locate_1d(long num, long den, long num_dx, long den_dx, int t, int* out) {
  long n, d;
  if (t) { n = num * den_dx; d = den * num_dx; }
  else { n = 2 * num * den_dx - den * num_dx; d = 2 * den * num_dx; }
  if (d == 0) return TW_EINVAL;
  if (n % d != 0) return TW_ERANGE;
  *out = (int)(n / d);
  return 0;
}

int or_locate(const tw_config* c, int grid, long num_x, long den_x, long num_y, long den_y, int* ix, int* iy) {
  if (c->den_dx <= 0 || c->num_dx <= 0) return TW_EINVAL;
  int rx, ry;
  int e = locate_1d(num_x, den_x, c->num_dx, c->den_dx, grid_tx(grid), &rx);
  if (e) return e;
  e = locate_1d(num_y, den_y, c->num_dx, c->den_dx, grid_ty(grid), &ry);
  if (e) return e;
  int rows = c->mx + grid_tx(grid), cols = c->my + grid_ty(grid);
  if (rx < 0 || rx >= rows || ry < 0 || ry >= cols) return TW_ERANGE;
  *ix = rx;
  *iy = ry;
  return 0;
}

/* Class_Grid::interpolate_forward_parameter, cuda/src/cpp/grid_parameter.cpp:6-60.
 * model: (nxm+1) x (nym+1) nodes, unpadded; the reference's extra padding row
 * and column only ever receive weight 0 and are read here as 0. */
int or_interpolate(const tw_config* c, int grid, const double* model, int nxm, int nym, long num_dxm, long den_dxm,
                   double* out) {
  int rows = c->mx + grid_tx(grid), cols = c->my + grid_ty(grid);
  long Nx = nxm + 1, Ny = nym + 1;
  long nd = c->num_dx, dd = c->den_dx;
  if (nd <= 0 || dd <= 0 || num_dxm <= 0 || den_dxm <= 0) return TW_EINVAL;
  for (int ix = 0; ix < rows; ix++)
    for (int iy = 0; iy < cols; iy++) {
      long nxl, dxl, nyl, dyl;
      if (grid_tx(grid)) { nxl = ix * nd; dxl = dd; } else { nxl = (2L * ix + 1) * nd; dxl = 2 * dd; }
      if (grid_ty(grid)) { nyl = iy * nd; dyl = dd; } else { nyl = (2L * iy + 1) * nd; dyl = 2 * dd; }
      long ixm = (nxl * den_dxm) / (dxl * num_dxm);
      long iym = (nyl * den_dxm) / (dyl * num_dxm);
      double xi = (double)(nxl * den_dxm) / (double)(dxl * num_dxm) - ixm;
      double eta = (double)(nyl * den_dxm) / (double)(dyl * num_dxm) - iym;
      if (ixm < 0 || ixm + 1 > Nx || iym < 0 || iym + 1 > Ny) { or_seterr("model too small for the grid"); return TW_ERANGE; }
#define MV(a, b) (((a) < Nx && (b) < Ny) ? model[(a) * Ny + (b)] : 0.0)
      double v = (1 - xi) * (1 - eta) * MV(ixm, iym) + (1 - xi) * (eta)*MV(ixm, iym + 1) +
                 (xi) * (1 - eta) * MV(ixm + 1, iym) + (xi) * (eta)*MV(ixm + 1, iym + 1);
#undef MV
      out[(size_t)ix * cols + iy] = v;
    }
  return 0;
}

/* forward_simulation.cpp:62-70 (velocity grids) and 104-112 (stress grids) */
double or_ricker_increment(const tw_config* c, double f0, double t0, int it, int half) {
  double dt = c->dt, dx = c->dx;
  double t = half ? (it + 1. / 2.) * (double)dt - t0 : it * (double)dt - t0;
  double A = (1 - 2 * M_PI * M_PI * f0 * f0 * t * t) * exp(-M_PI * M_PI * f0 * f0 * t * t);
  return A * dt / (dx * dx) / c->source_scaling;
}

/* round a coefficient value to its storage type */
static double round_coef(const tw_config* c, double v) {
  switch (c->precision) {
    case TW_FP64: return v;
    case TW_FP32: return (double)(float)v;
    case TW_BF16: return c->coef16 ? (double)bf16_to_f32(f32_to_bf16((float)v)) : (double)(float)v;
    default: return c->coef16 ? (double)(_Float16)(float)v : (double)(float)v;
  }
}

static int coef_used(int physics, int k) {
  if (physics == TW_ELASTIC) return k != TW_C_KAPPA;
  return k == TW_C_BINV_VX || k == TW_C_BINV_VY || k == TW_C_KAPPA;
}

/* impedances Z = sqrt(M/binv) = rho*c along the sides of each velocity field */
static void compute_z(or_solver* o) {
  or_common* cm = &o->cm;
  const tw_config* c = &cm->cfg;
  int anyc = 0;
  for (int k = 0; k < 4; k++) anyc |= c->bc[k] == 'C';
  if (!anyc) return;
  int el = c->physics == TW_ELASTIC;
  int Mx = c->mx, My = c->my;
  for (int side = 0; side < 4; side++) {
    if (c->bc[side] != 'C') continue;
    int axis = side / 2, right = side % 2;
    for (int f = TW_VX; f <= TW_VY; f++) {
      int normal = (axis == 0) == (f == TW_VX);
      if (!el && !normal) continue;
      int rows = cm->rows[f], cols = cm->cols[f];
      int n = axis == 0 ? cols : rows;
      free(cm->z[f][side]);
      cm->z[f][side] = (double*)malloc((size_t)n * sizeof(double));
      cm->zlen[f][side] = n;
      const double* binv = cm->coef[f == TW_VX ? TW_C_BINV_VX : TW_C_BINV_VY];
      for (int j = 0; j < n; j++) {
        int a = axis == 0 ? (right ? rows - 1 : 0) : (right ? cols - 1 : 0);
        size_t qv = axis == 0 ? (size_t)a * cols + j : (size_t)j * cols + a;
        double mod;
        if (normal) { /* normal velocity pairs with the normal stress on MM */
          int am = axis == 0 ? (right ? Mx - 1 : 0) : (right ? My - 1 : 0);
          size_t qm = axis == 0 ? (size_t)am * My + j : (size_t)j * My + am;
          mod = el ? cm->coef[TW_C_LAMBDA][qm] + cm->coef[TW_C_TWOMU][qm] : cm->coef[TW_C_KAPPA][qm];
        } else { /* tangential velocity pairs with sxy on NN */
          int an = axis == 0 ? (right ? Mx : 0) : (right ? My : 0);
          size_t qn = axis == 0 ? (size_t)an * (My + 1) + j : (size_t)j * (My + 1) + an;
          mod = cm->coef[TW_C_MU][qn];
        }
        cm->z[f][side][j] = sqrt(mod / binv[qv]);
      }
    }
  }
  o->vt->load_z(o->impl);
}

static void set_default_material(or_solver* o) {
  or_common* cm = &o->cm;
  const tw_config* c = &cm->cfg;
  double rho = 1., vp = c->physics == TW_ELASTIC ? 2. : 1., vs = 1.;
  for (int k = 0; k < 6; k++) {
    if (!cm->coef[k]) continue;
    double v;
    switch (k) {
      case TW_C_BINV_VX:
      case TW_C_BINV_VY: v = 1. / rho; break;
      case TW_C_MU: v = rho * vs * vs; break;
      case TW_C_TWOMU: v = 2 * rho * vs * vs; break;
      case TW_C_LAMBDA: v = rho * vp * vp - 2 * rho * vs * vs; break;
      default: v = rho * vp * vp; break;
    }
    v = round_coef(c, v);
    size_t n = (size_t)cm->coef_rows[k] * cm->coef_cols[k];
    for (size_t i = 0; i < n; i++) cm->coef[k][i] = v;
    o->vt->load_coef(o->impl, k);
  }
  compute_z(o);
}

or_solver* or_create(const tw_config* cfg) {
  if (cfg_check(cfg)) return NULL;
  or_solver* o = (or_solver*)calloc(1, sizeof(or_solver));
  if (!o) return NULL;
  or_common* cm = &o->cm;
  cm->cfg = *cfg;
  cm->src_grid = -1;
  if (cfg->physics == TW_ELASTIC) {
    int fl[5] = {TW_VX, TW_VY, TW_SXX, TW_SYY, TW_SXY};
    cm->nf = 5;
    memcpy(cm->fields, fl, sizeof fl);
  } else {
    int fl[3] = {TW_VX, TW_VY, TW_P};
    cm->nf = 3;
    memcpy(cm->fields, fl, sizeof fl);
  }
  for (int k = 0; k < cm->nf; k++) {
    int f = cm->fields[k];
    or_field_shape(cfg, f, &cm->rows[f], &cm->cols[f]);
  }
  for (int k = 0; k < 6; k++) {
    if (!coef_used(cfg->physics, k)) continue;
    int g = coef_grid(k);
    cm->coef_rows[k] = cfg->mx + grid_tx(g);
    cm->coef_cols[k] = cfg->my + grid_ty(g);
    cm->coef[k] = (double*)calloc((size_t)cm->coef_rows[k] * cm->coef_cols[k], sizeof(double));
  }
  int nt = cfg->nt_max > 0 ? cfg->nt_max : 1;
  cm->ek = (double*)calloc((size_t)nt, sizeof(double));
  cm->ep0 = (double*)calloc((size_t)nt, sizeof(double));
  cm->ep1 = (double*)calloc((size_t)nt, sizeof(double));
  switch (cfg->precision) {
    case TW_FP64: o->vt = &f64_vt; break;
    case TW_FP32: o->vt = &f32_vt; break;
    case TW_BF16: o->vt = &bf16_vt; break;
    default: o->vt = &f16_vt; break;
  }
  o->impl = o->vt->create(cm);
  if (!o->impl) { or_seterr("allocation failed"); or_destroy(o); return NULL; }
  set_default_material(o);
  return o;
}

void or_destroy(or_solver* o) {
  if (!o) return;
  or_common* cm = &o->cm;
  if (o->impl) o->vt->destroy(o->impl);
  for (int k = 0; k < 6; k++) {
    free(cm->coef[k]);
    for (int s = 0; s < 4; s++) free(cm->z[k][s]);
  }
  free(cm->src_inc);
  free(cm->rf); free(cm->rix); free(cm->riy); free(cm->rec);
  free(cm->ek); free(cm->ep0); free(cm->ep1);
  free(o);
}

int or_set_coef(or_solver* o, int k, const double* v) {
  or_common* cm = &o->cm;
  if (k < 0 || k > 5 || !cm->coef[k]) { or_seterr("coefficient not used by this physics"); return TW_EINVAL; }
  size_t n = (size_t)cm->coef_rows[k] * cm->coef_cols[k];
  for (size_t i = 0; i < n; i++) cm->coef[k][i] = round_coef(&cm->cfg, v[i]);
  o->vt->load_coef(o->impl, k);
  compute_z(o);
  cm->ep_valid = 0;
  return 0;
}

int or_get_coef(or_solver* o, int k, double* v) {
  or_common* cm = &o->cm;
  if (k < 0 || k > 5 || !cm->coef[k]) return TW_EINVAL;
  memcpy(v, cm->coef[k], (size_t)cm->coef_rows[k] * cm->coef_cols[k] * sizeof(double));
  return 0;
}

/* retrieve_forward_parameter + make_density_reciprocal
 * (cuda/src/cpp/grid_parameter.cpp:75-162): lambda = rho vp^2 - 2 mu,
 * 2 mu = 2 rho vs^2, mu = rho vs^2 formed on the model nodes, then
 * interpolated; 1/rho is 1 / (interpolated rho). */
int or_set_model(or_solver* o, const double* rho, const double* vp, const double* vs, int nxm, int nym, long num_dxm,
                 long den_dxm) {
  or_common* cm = &o->cm;
  const tw_config* c = &cm->cfg;
  size_t nm = (size_t)(nxm + 1) * (nym + 1);
  double* a = (double*)malloc(nm * sizeof(double));
  double* b = (double*)malloc(nm * sizeof(double));
  int e = 0;
  for (int k = 0; k < 6 && !e; k++) {
    if (!cm->coef[k]) continue;
    int g = coef_grid(k);
    size_t n = (size_t)cm->coef_rows[k] * cm->coef_cols[k];
    double* out = (double*)malloc(n * sizeof(double));
    switch (k) {
      case TW_C_BINV_VX:
      case TW_C_BINV_VY:
        e = or_interpolate(c, g, rho, nxm, nym, num_dxm, den_dxm, out);
        for (size_t i = 0; i < n; i++) out[i] = 1. / out[i];
        break;
      case TW_C_MU:
        for (size_t i = 0; i < nm; i++) a[i] = rho[i] * vs[i] * vs[i];
        e = or_interpolate(c, g, a, nxm, nym, num_dxm, den_dxm, out);
        break;
      case TW_C_TWOMU:
        for (size_t i = 0; i < nm; i++) a[i] = 2 * rho[i] * vs[i] * vs[i];
        e = or_interpolate(c, g, a, nxm, nym, num_dxm, den_dxm, out);
        break;
      case TW_C_LAMBDA:
        for (size_t i = 0; i < nm; i++) {
          b[i] = 2 * rho[i] * vs[i] * vs[i];
          a[i] = rho[i] * vp[i] * vp[i] - b[i];
        }
        e = or_interpolate(c, g, a, nxm, nym, num_dxm, den_dxm, out);
        break;
      default:
        for (size_t i = 0; i < nm; i++) a[i] = rho[i] * vp[i] * vp[i];
        e = or_interpolate(c, g, a, nxm, nym, num_dxm, den_dxm, out);
        break;
    }
    if (!e) {
      for (size_t i = 0; i < n; i++) cm->coef[k][i] = round_coef(c, out[i]);
      o->vt->load_coef(o->impl, k);
    }
    free(out);
  }
  free(a);
  free(b);
  if (!e) compute_z(o);
  cm->ep_valid = 0;
  return e;
}

static int check_point(or_common* cm, int grid, int ix, int iy, int source) {
  const tw_config* c = &cm->cfg;
  int rows = c->mx + grid_tx(grid), cols = c->my + grid_ty(grid);
  if (ix < 0 || ix >= rows || iy < 0 || iy >= cols) return TW_ERANGE;
  if (source) {
    /* forward.cpp:95-98: a source on a weak (SBP) node boundary is rejected */
    if (c->bc[0] != 'P' && grid_tx(grid) && (ix == 0 || ix == rows - 1)) return TW_ERANGE;
    if (c->bc[2] != 'P' && grid_ty(grid) && (iy == 0 || iy == cols - 1)) return TW_ERANGE;
  }
  return 0;
}

int or_set_source(or_solver* o, int grid, int ix, int iy, const double* inc, int n) {
  or_common* cm = &o->cm;
  if (grid < 0 || grid > 3) return TW_EINVAL;
  int e = check_point(cm, grid, ix, iy, 1);
  if (e) return e;
  free(cm->src_inc);
  cm->src_inc = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
  memcpy(cm->src_inc, inc, (size_t)n * sizeof(double));
  cm->src_n = n;
  cm->src_grid = grid;
  cm->src_ix = ix;
  cm->src_iy = iy;
  return 0;
}

int or_set_source_ricker(or_solver* o, int grid, int ix, int iy, double f0, double t0) {
  or_common* cm = &o->cm;
  int n = cm->cfg.nt_max;
  int half = !(grid == TW_GRID_MN || grid == TW_GRID_NM);
  double* inc = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
  for (int it = 0; it < n; it++) inc[it] = or_ricker_increment(&cm->cfg, f0, t0, it, half);
  int e = or_set_source(o, grid, ix, iy, inc, n);
  free(inc);
  return e;
}

int or_set_receivers(or_solver* o, int n, const int* field, const int* ix, const int* iy) {
  or_common* cm = &o->cm;
  for (int r = 0; r < n; r++) {
    int rows, cols;
    if (or_field_shape(&cm->cfg, field[r], &rows, &cols)) return TW_EINVAL;
    if (ix[r] < 0 || ix[r] >= rows || iy[r] < 0 || iy[r] >= cols) return TW_ERANGE;
  }
  free(cm->rf); free(cm->rix); free(cm->riy); free(cm->rec);
  cm->nrcv = n;
  size_t m = (size_t)(n > 0 ? n : 1);
  cm->rf = (int*)malloc(m * sizeof(int));
  cm->rix = (int*)malloc(m * sizeof(int));
  cm->riy = (int*)malloc(m * sizeof(int));
  cm->rec = (double*)calloc(m * (size_t)(cm->cfg.nt_max > 0 ? cm->cfg.nt_max : 1), sizeof(double));
  memcpy(cm->rf, field, (size_t)n * sizeof(int));
  memcpy(cm->rix, ix, (size_t)n * sizeof(int));
  memcpy(cm->riy, iy, (size_t)n * sizeof(int));
  return 0;
}

int or_step(or_solver* o, int nsteps) {
  for (int k = 0; k < nsteps; k++) o->vt->step(o->impl);
  return 0;
}

int or_read_receivers(or_solver* o, double* out, int* nt) {
  or_common* cm = &o->cm;
  int n = cm->it < cm->cfg.nt_max ? cm->it : cm->cfg.nt_max;
  *nt = n;
  if (out)
    for (int r = 0; r < cm->nrcv; r++)
      for (int t = 0; t < n; t++)
        out[(size_t)r * n + t] = cm->cfg.source_scaling * cm->rec[(size_t)r * cm->cfg.nt_max + t];
  return 0;
}

int or_read_energy(or_solver* o, double* ek, double* ep0, double* ep1, int* nt) {
  or_common* cm = &o->cm;
  int n = cm->it < cm->cfg.nt_max ? cm->it : cm->cfg.nt_max;
  double s2 = cm->cfg.source_scaling * cm->cfg.source_scaling;
  *nt = n;
  for (int t = 0; t < n; t++) {
    if (ek) ek[t] = s2 * cm->ek[t];
    if (ep0) ep0[t] = s2 * cm->ep0[t];
    if (ep1) ep1[t] = s2 * cm->ep1[t];
  }
  return 0;
}

int or_get_field(or_solver* o, int f, double* out) {
  int r, c;
  if (or_field_shape(&o->cm.cfg, f, &r, &c)) return TW_EINVAL;
  o->vt->get_field(o->impl, f, out, 0);
  return 0;
}

int or_set_field(or_solver* o, int f, const double* in) {
  int r, c;
  if (or_field_shape(&o->cm.cfg, f, &r, &c)) return TW_EINVAL;
  o->vt->set_field(o->impl, f, in, 0);
  o->cm.ep_valid = 0;
  return 0;
}

int or_get_companion(or_solver* o, int f, double* out) {
  int r, c;
  if (or_field_shape(&o->cm.cfg, f, &r, &c)) return TW_EINVAL;
  int cp = o->cm.cfg.cpst;
  int which = (cp == TW_CPST_3C || cp == TW_CPST_M3C || cp == TW_CPST_6C) ? 2 : 1;
  o->vt->get_field(o->impl, f, out, which);
  return 0;
}

int or_table(or_solver* o, int which, double* out) { return o->vt->table(o->impl, which, out); }

/* ---- single-point update and residual-code entry points (tests only) ---- */
int or_update_points(int precision, int cpst, long n, double* s, double* r, double* c) {
  if (cpst < TW_CPST_0 || cpst > TW_CPST_6C) return TW_EINVAL;
  if (precision == TW_FP64) f64_update_points(cpst, n, s, r, c);
  else if (precision == TW_FP32) f32_update_points(cpst, n, s, r, c);
  else return TW_EINVAL;
  return 0;
}

int or_store16_points(int precision, int cpst, long n, const double* t, double* hi, double* res) {
  if (cpst != TW_CPST_0 && cpst != TW_CPST_K && !OR_K8(cpst)) return TW_EINVAL;
  if (precision == TW_BF16) bf16_store_points(cpst, n, t, hi, res);
  else if (precision == TW_FP16) f16_store_points(cpst, n, t, hi, res);
  else return TW_EINVAL;
  return 0;
}

int or_k8_decode_points(int precision, long n, const double* hi, const double* q, double* out) {
  if (precision == TW_BF16) bf16_k8_decode_points(n, hi, q, out);
  else if (precision == TW_FP16) f16_k8_decode_points(n, hi, q, out);
  else return TW_EINVAL;
  return 0;
}
