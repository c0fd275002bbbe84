// tw_host.cpp -- host side of the C-ABI (include/tw.h).
//
// Setup runs in FP64 on the host exactly as the reference builds its grids
// (cuda/src/cpp/grid.hpp:146-481, grid_parameter.cpp, grid_energy.cpp,
// forward.cpp); the time loop launches the fused gfx950 kernels of
// tw_kernels.hip on one HIP stream with no host synchronisation inside
// tw_step.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <stdlib.h>

#include <algorithm>
#include <utility>
#include <array>
#include <map>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/tw.h"
#include "tw_internal.h"

using tw::Plan;

namespace {

// ---- operator constants: cuda/src/cpp/namespace_forward.hpp:65-156 ----
const double kStencil[4] = {1. / 24, -9. / 8, 9. / 8, -1. / 24};
const double kDM_L[4][5] = {{-2., 3., -1., 0., 0.},
                            {-1., 1., 0., 0., 0.},
                            {1. / 24., -9. / 8., 9. / 8., -1. / 24., 0.},
                            {-1. / 71., 6. / 71., -83. / 71., 81. / 71., -3. / 71.}};
const double kDM_R[4][5] = {{3. / 71., -81. / 71., 83. / 71., -6. / 71., 1. / 71.},
                            {0., 1. / 24., -9. / 8., 9. / 8., -1. / 24.},
                            {0., 0., 0., -1., 1.},
                            {0., 0., 1., -3., 2.}};
const double kDN_L[3][5] = {{-79. / 78., 27. / 26., -1. / 26., 1. / 78., 0.},
                            {2. / 21., -9. / 7., 9. / 7., -2. / 21., 0.},
                            {1. / 75., 0., -27. / 25., 83. / 75., -1. / 25.}};
const double kDN_R[3][5] = {{1. / 25., -83. / 75., 27. / 25., 0., -1. / 75.},
                            {0., 2. / 21., -9. / 7., 9. / 7., -2. / 21.},
                            {0., -1. / 78., 1. / 26., -27. / 26., 79. / 78.}};
const double kAN_L[4] = {7. / 18., 9. / 8., 1., 71. / 72.};
const double kAN_R[4] = {71. / 72., 1., 9. / 8., 7. / 18.};
const double kAM_L[3] = {13. / 12., 7. / 8., 25. / 24.};
const double kAM_R[3] = {25. / 24., 7. / 8., 13. / 12.};
const double kPM_L[3] = {15. / 8., -5. / 4., 3. / 8.};
const double kPM_R[3] = {3. / 8., -5. / 4., 15. / 8.};

thread_local std::string g_err;

int fail(std::string* where, int code, const std::string& msg) {
  if (where) *where = msg;
  g_err = msg;
  return code;
}

int field_grid(int f) {
  switch (f) {
    case TW_VX: return TW_GRID_NM;
    case TW_VY: return TW_GRID_MN;
    case TW_SXY: return TW_GRID_NN;
    default: return TW_GRID_MM;
  }
}
int grid_tx(int g) { return g == TW_GRID_NN || g == TW_GRID_NM; }
int grid_ty(int g) { return g == TW_GRID_NN || g == TW_GRID_MN; }
int coef_grid(int k) {
  switch (k) {
    case TW_C_BINV_VX: return TW_GRID_NM;
    case TW_C_BINV_VY: return TW_GRID_MN;
    case TW_C_MU: return TW_GRID_NN;
    default: return TW_GRID_MM;
  }
}
bool field_in(int physics, int f) {
  if (f < 0 || f > 5) return false;
  if (physics == TW_ELASTIC) return f != TW_P;
  return f == TW_VX || f == TW_VY || f == TW_P;
}
bool coef_in(int physics, int k) {
  if (k < 0 || k > 5) return false;
  if (physics == TW_ELASTIC) return k != TW_C_KAPPA;
  return k == TW_C_BINV_VX || k == TW_C_BINV_VY || k == TW_C_KAPPA;
}
bool is_velocity(int f) { return f == TW_VX || f == TW_VY; }
bool ref_precision(int p) { return p == TW_FP64 || p == TW_FP32; }
bool k8mode(int c) { return c == TW_CPST_K8 || c == TW_CPST_K8S; }
bool rmode(int c) { return c == TW_CPST_3R || c == TW_CPST_M3R || c == TW_CPST_6R || c == TW_CPST_K || k8mode(c); }
bool cmode(int c) { return c == TW_CPST_3C || c == TW_CPST_M3C || c == TW_CPST_6C; }

// ---- host conversions (round to nearest even) ----
uint16_t f32_to_bf16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
float bf16_to_f32(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

size_t state_size(int precision) { return precision == TW_FP64 ? 8 : (precision == TW_FP32 ? 4 : 2); }
// element size of array `which` of a field (0 state, 1 R companion, 2 C
// companion): the K8 residual holds one byte code per value
size_t array_size(const tw_config* c, int which) {
  return which == 1 && k8mode(c->cpst) ? 1 : state_size(c->precision);
}
size_t coef_size(int precision, int coef16) {
  if (precision == TW_FP64) return 8;
  if (precision == TW_FP32) return 4;
  return coef16 == 2 ? 1 : (coef16 ? 2 : 4);  // 3: 16-bit codes
}
size_t arith_size(int precision) { return precision == TW_FP64 ? 8 : 4; }

// double -> storage bytes, and back
void to_storage(int precision, const double* in, void* out, size_t n) {
  switch (precision) {
    case TW_FP64: memcpy(out, in, n * 8); break;
    case TW_FP32:
      for (size_t i = 0; i < n; i++) ((float*)out)[i] = (float)in[i];
      break;
    case TW_BF16:
      for (size_t i = 0; i < n; i++) ((uint16_t*)out)[i] = f32_to_bf16((float)in[i]);
      break;
    default:
      for (size_t i = 0; i < n; i++) ((_Float16*)out)[i] = (_Float16)(float)in[i];
      break;
  }
}
void from_storage(int precision, const void* in, double* out, size_t n) {
  switch (precision) {
    case TW_FP64: memcpy(out, in, n * 8); break;
    case TW_FP32:
      for (size_t i = 0; i < n; i++) out[i] = ((const float*)in)[i];
      break;
    case TW_BF16:
      for (size_t i = 0; i < n; i++) out[i] = bf16_to_f32(((const uint16_t*)in)[i]);
      break;
    default:
      for (size_t i = 0; i < n; i++) out[i] = (double)(float)((const _Float16*)in)[i];
      break;
  }
}
// coefficient rounding: its storage type (the reference casts to the solver precision)
double round_coef(const tw_config* c, double v) {
  if (c->coef16 >= 2 && !ref_precision(c->precision)) return v;  // quantized per row at upload
  switch (c->precision) {
    case TW_FP64: return v;
    case TW_FP32: return (double)(float)v;
    case TW_BF16: return c->coef16 ? (double)bf16_to_f32(f32_to_bf16((float)v)) : (double)(float)v;
    default: return c->coef16 ? (double)(float)(_Float16)(float)v : (double)(float)v;
  }
}
int coef_store_precision(const tw_config* c) {
  if (c->precision == TW_FP64) return TW_FP64;
  if (c->precision == TW_FP32 || !c->coef16) return TW_FP32;
  return c->precision;
}
// a value rounded to the arithmetic type, returned as double
double arith(int precision, double v) { return precision == TW_FP64 ? v : (double)(float)v; }

int check_config(const tw_config* c, std::string* err) {
  if (!c) return fail(err, TW_EINVAL, "null config");
  if (c->physics != TW_ELASTIC && c->physics != TW_ACOUSTIC) return fail(err, TW_EINVAL, "physics must be TW_ELASTIC or TW_ACOUSTIC");
  if (c->mx < 8 || c->my < 8) return fail(err, TW_EINVAL, "mx and my must be >= 8");
  if ((int64_t)c->mx * c->my > (int64_t)1 << 34) return fail(err, TW_ENOMEM, "grid too large");
  if (!(c->dx > 0) || !(c->dt > 0)) return fail(err, TW_EINVAL, "dx and dt must be > 0");
  for (int k = 0; k < 4; k++)
    if (c->bc[k] != 'P' && c->bc[k] != 'F' && c->bc[k] != 'C') return fail(err, TW_EINVAL, "bc must be 'P', 'F' or 'C'");
  for (int a = 0; a < 2; a++)
    if ((c->bc[2 * a] == 'P') != (c->bc[2 * a + 1] == 'P')) return fail(err, TW_EINVAL, "periodic sides must pair");
  if (c->precision < TW_FP64 || c->precision > TW_FP16) return fail(err, TW_EINVAL, "precision");
  if (c->cpst < TW_CPST_0 || c->cpst > TW_CPST_K8S) return fail(err, TW_EINVAL, "cpst");
  const bool k16 = c->cpst == TW_CPST_K || k8mode(c->cpst);
  if (ref_precision(c->precision) && k16) return fail(err, TW_EINVAL, "cpst K and K8 need 16-bit storage");
  if (!ref_precision(c->precision) && c->cpst != TW_CPST_0 && !k16)
    return fail(err, TW_EINVAL, "16-bit storage supports cpst 0, K or K8");
  if (c->energy_form == 1 && (c->physics != TW_ACOUSTIC || c->bc[0] != 'P' || c->bc[2] != 'P'))
    return fail(err, TW_EINVAL, "energy_form 1 is the periodic acoustic reference form");
  if (c->energy_form != 0 && c->energy_form != 1) return fail(err, TW_EINVAL, "energy_form");
  if (c->coef16 < 0 || c->coef16 > 3) return fail(err, TW_EINVAL, "coef16 must be 0, 1, 2 or 3");
  if (c->coef16 >= 2 && ref_precision(c->precision))
    return fail(err, TW_EINVAL, "coefficient codes (coef16 2, 3) need 16-bit storage");
  if (c->nt_max < 0) return fail(err, TW_EINVAL, "nt_max must be >= 0");
  if (!(c->source_scaling > 0)) return fail(err, TW_EINVAL, "source_scaling must be > 0");
  return TW_OK;
}

void shape_of(const tw_config* c, int g, int* rows, int* cols) {
  *rows = c->mx + grid_tx(g);
  *cols = c->my + grid_ty(g);
}

// Operator tables (set_forward_operators, cuda/src/cpp/grid.hpp:274-464):
// built in FP64, cast to the arithmetic type; SAT lifts P/(dx*A) are cast,
// then multiplied by dt (itself cast first) and cast again
// (container_run_time.hpp:124).
void build_tables(const tw_config* c, tw::Tables* t) {
  const int P = c->precision;
  const double dt = c->dt, dx = c->dx;
  for (int k = 0; k < 4; k++) t->st[k] = arith(P, c->stencil_form ? kStencil[k] * dt / dx : kStencil[k] * (dt / dx));
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 5; j++) {
      t->DN_L[i][j] = arith(P, kDM_L[i][j] * (dt / dx));
      t->DN_R[i][j] = arith(P, kDM_R[i][j] * (dt / dx));
    }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 5; j++) {
      t->DM_L[i][j] = arith(P, kDN_L[i][j] * (dt / dx));
      t->DM_R[i][j] = arith(P, kDN_R[i][j] * (dt / dx));
    }
  const double dtc = arith(P, dt);
  t->ainvN_L = arith(P, arith(P, 1.0 / (dx * kAN_L[0])) * dtc);
  t->ainvN_R = arith(P, arith(P, 1.0 / (dx * kAN_R[3])) * dtc);
  for (int i = 0; i < 3; i++) {
    t->ainvM_L[i] = arith(P, arith(P, kPM_L[i] / (dx * kAM_L[i])) * dtc);
    t->ainvM_R[i] = arith(P, arith(P, kPM_R[i] / (dx * kAM_R[i])) * dtc);
    t->projM_L[i] = arith(P, kPM_L[i]);
    t->projM_R[i] = arith(P, kPM_R[i]);
  }
  for (int i = 0; i < 4; i++) t->hN[i] = arith(P, kAN_L[i] * dx);
  for (int i = 0; i < 3; i++) t->hM[i] = arith(P, kAM_L[i] * dx);
}

// cuda/src/cpp/forward.cpp:30-54
int locate_1d(int64_t num, int64_t den, int64_t num_dx, int64_t den_dx, int node, int32_t* out) {
  int64_t n, d;
  if (node) {
    n = num * den_dx;
    d = den * num_dx;
  } else {
    n = 2 * num * den_dx - den * num_dx;
    d = 2 * den * num_dx;
  }
  if (d == 0) return TW_EINVAL;
  if (n % d != 0) return TW_ERANGE;
  *out = (int32_t)(n / d);
  return TW_OK;
}

// cuda/src/cpp/grid_parameter.cpp:6-60 (bilinear, fixed term order)
// Model node (ixm, iym) and bilinear weights (xi, eta) of point (ix, iy) of
// grid g; false when the model does not cover the point.
struct InterpAt {
  int64_t ixm, iym;
  double xi, eta;
};
bool interp_at(const tw_config* c, int g, int ix, int iy, int64_t num_dxm, int64_t den_dxm, int64_t Nx, int64_t Ny,
               InterpAt* w) {
  const int64_t nd = c->num_dx, dd = c->den_dx;
  int64_t nxl, dxl, nyl, dyl;
  if (grid_tx(g)) { nxl = ix * nd; dxl = dd; } else { nxl = (2 * (int64_t)ix + 1) * nd; dxl = 2 * dd; }
  if (grid_ty(g)) { nyl = iy * nd; dyl = dd; } else { nyl = (2 * (int64_t)iy + 1) * nd; dyl = 2 * dd; }
  w->ixm = (nxl * den_dxm) / (dxl * num_dxm);
  w->iym = (nyl * den_dxm) / (dyl * num_dxm);
  w->xi = (double)(nxl * den_dxm) / (double)(dxl * num_dxm) - w->ixm;
  w->eta = (double)(nyl * den_dxm) / (double)(dyl * num_dxm) - w->iym;
  return !(w->ixm < 0 || w->ixm + 1 > Nx || w->iym < 0 || w->iym + 1 > Ny);
}

int interp_check(const tw_config* c, int32_t nxm, int32_t nym, int64_t num_dxm, int64_t den_dxm) {
  if (c->num_dx <= 0 || c->den_dx <= 0)
    return fail(nullptr, TW_EINVAL, "tw_config num_dx/den_dx must be set for model interpolation");
  if (num_dxm <= 0 || den_dxm <= 0 || nxm < 1 || nym < 1) return fail(nullptr, TW_EINVAL, "model spacing/size");
  return TW_OK;
}

int interpolate(const tw_config* c, int g, const double* model, int32_t nxm, int32_t nym, int64_t num_dxm,
                int64_t den_dxm, double* out) {
  int rows, cols;
  shape_of(c, g, &rows, &cols);
  const int64_t Nx = (int64_t)nxm + 1, Ny = (int64_t)nym + 1;
  int e = interp_check(c, nxm, nym, num_dxm, den_dxm);
  if (e) return e;
  auto mv = [&](int64_t a, int64_t b) { return (a < Nx && b < Ny) ? model[a * Ny + b] : 0.0; };
  for (int ix = 0; ix < rows; ix++)
    for (int iy = 0; iy < cols; iy++) {
      InterpAt w;
      if (!interp_at(c, g, ix, iy, num_dxm, den_dxm, Nx, Ny, &w))
        return fail(nullptr, TW_ERANGE, "model does not cover the simulation grid");
      const int64_t ixm = w.ixm, iym = w.iym;
      const double xi = w.xi, eta = w.eta;
      const double v = (1 - xi) * (1 - eta) * mv(ixm, iym) + (1 - xi) * (eta)*mv(ixm, iym + 1) +
                       (xi) * (1 - eta) * mv(ixm + 1, iym) + (xi) * (eta)*mv(ixm + 1, iym + 1);
      out[(size_t)ix * cols + iy] = v;
    }
  return TW_OK;
}

// transpose of interpolate: out (model nodes) = sum over the grid's points of
// each weight times the point's value (nodes beyond the model read as 0 by
// interpolate receive nothing)
int interpolate_adjoint(const tw_config* c, int g, const double* vals, int32_t nxm, int32_t nym, int64_t num_dxm,
                        int64_t den_dxm, double* out) {
  int rows, cols;
  shape_of(c, g, &rows, &cols);
  const int64_t Nx = (int64_t)nxm + 1, Ny = (int64_t)nym + 1;
  int e = interp_check(c, nxm, nym, num_dxm, den_dxm);
  if (e) return e;
  std::fill(out, out + Nx * Ny, 0.0);
  auto add = [&](int64_t a, int64_t b, double v) {
    if (a < Nx && b < Ny) out[a * Ny + b] += v;
  };
  for (int ix = 0; ix < rows; ix++)
    for (int iy = 0; iy < cols; iy++) {
      InterpAt w;
      if (!interp_at(c, g, ix, iy, num_dxm, den_dxm, Nx, Ny, &w))
        return fail(nullptr, TW_ERANGE, "model does not cover the simulation grid");
      const double y = vals[(size_t)ix * cols + iy];
      add(w.ixm, w.iym, (1 - w.xi) * (1 - w.eta) * y);
      add(w.ixm, w.iym + 1, (1 - w.xi) * (w.eta) * y);
      add(w.ixm + 1, w.iym, (w.xi) * (1 - w.eta) * y);
      add(w.ixm + 1, w.iym + 1, (w.xi) * (w.eta) * y);
    }
  return TW_OK;
}

// forward_simulation.cpp:62-70 / 104-112; device driver divides by source_scaling
double ricker_inc(const tw_config* c, double f0, double t0, int32_t it, int32_t half) {
  const double dt = c->dt, dx = c->dx;
  const double t = half ? (it + 1. / 2.) * (double)dt - t0 : it * (double)dt - t0;
  const double A = (1 - 2 * M_PI * M_PI * f0 * f0 * t * t) * exp(-M_PI * M_PI * f0 * f0 * t * t);
  return A * dt / (dx * dx) / c->source_scaling;
}

}  // namespace

// ---------------------------------------------------------------- solver
struct tw_solver {
  tw_config cfg{};
  Plan plan{};
  std::string err;
  std::vector<double> hcoef[6];
  std::vector<void*> allocs;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int src_grid = -1, src_ix = 0, src_iy = 0;
  std::vector<double> src_inc;
  std::vector<tw::Receiver> rcv_v, rcv_s;
  int nrcv = 0;
  size_t part_cap = 0;  // doubles in each of plan.part_v / part_s
  int it = 0;
  // bumped by every change that a batch's captured tables depend on (source,
  // receivers, coefficients): a batch rebuilds its graph when it differs
  uint64_t gen = 0;
  bool ep_valid = false;
  // steps recomputed inside the adjoint sweep: their energies and receiver
  // values were recorded by the forward run, so these steps record nothing
  bool quiet = false;
  bool need_uhat = false;
  // small grid: single steps run the per-point kernels (tw_create's choice,
  // not a TW_KERNEL request); batches still run the streaming kernels
  bool auto_point = false;
  // discrete adjoint: checkpoints of the last checkpointed forward run
  tw::AdjPlan adj{};
  bool adj_taps = false;
  void* adj_ckpt = nullptr;  // [nckpt] states
  void* adj_seg = nullptr;   // [interval + 1] states
  void* adj_final = nullptr; // state after the run
  double* adj_inj = nullptr;
  size_t adj_state_bytes = 0, adj_inj_n = 0;
  int adj_it0 = 0, adj_n = 0, adj_I = 0, adj_nckpt = 0, adj_cap_ckpt = 0, adj_cap_seg = 0;
  bool adj_done = false;     // a checkpointed run is available
  // per-launch timing: 4 events per profiled step
  std::vector<hipEvent_t> prof_ev;
  int prof_n = 0, prof_max = 0;
};

namespace {

int hip_fail(tw_solver* s, hipError_t e, const char* what) {
  return fail(s ? &s->err : nullptr, TW_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(s, call)                                 \
  do {                                                  \
    hipError_t e_ = (call);                             \
    if (e_ != hipSuccess) return hip_fail((s), e_, #call); \
  } while (0)

int dalloc(tw_solver* s, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes > 0 ? bytes : 256);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(&s->err, TW_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  s->allocs.push_back(*p);
  e = hipMemset(*p, 0, bytes > 0 ? bytes : 256);
  if (e != hipSuccess) return hip_fail(s, e, "hipMemset");
  return TW_OK;
}

void dfree(tw_solver* s, void** p) {
  if (!*p) return;
  (void)hipFree(*p);
  for (auto& a : s->allocs)
    if (a == *p) a = nullptr;
  *p = nullptr;
}

// Coefficient codes: each row is mapped affinely onto [0, levels] (levels
// 255 for 8-bit, 65535 for 16-bit codes; offset = the row minimum, step =
// range / levels, both FP32), every value takes the code whose FP32 decode
// off + q * step is nearest, and the host copy is replaced by the decoded
// values, which is what the kernels compute with.
template <class Q>
void quantize_row(double* v, int n, Q* q, float* off, float* step, int levels) {
  double lo = v[0], hi = v[0];
  for (int i = 1; i < n; i++) {
    lo = std::min(lo, v[i]);
    hi = std::max(hi, v[i]);
  }
  const float o = (float)lo;
  const float st = (float)((hi - lo) / (double)levels);
  for (int i = 0; i < n; i++) {
    int c0 = st > 0 ? (int)std::lround((v[i] - (double)o) / (double)st) : 0;
    int best = 0;
    double err = 0.0;
    for (int c = c0 - 1; c <= c0 + 1; c++) {
      const int cc = std::min(levels, std::max(0, c));
      const double e = std::fabs((double)tw::c8_dec((unsigned)cc, o, st) - v[i]);
      if (c == c0 - 1 || e < err) {
        best = cc;
        err = e;
      }
    }
    q[i] = (Q)best;
    v[i] = (double)tw::c8_dec((unsigned)best, o, st);
  }
  *off = o;
  *step = st;
}

int upload_coef(tw_solver* s, int k) {
  const tw_config* c = &s->cfg;
  tw::DevCoef& d = s->plan.k[k];
  if (c->coef16 == 2 && !ref_precision(c->precision)) {
    std::vector<uint8_t> codes((size_t)d.rows * d.pitch, 0);
    std::vector<float> rp((size_t)d.rows * 2);
    for (int r = 0; r < d.rows; r++)
      quantize_row(s->hcoef[k].data() + (size_t)r * d.cols, d.cols, codes.data() + (size_t)r * d.pitch, &rp[2 * r],
                   &rp[2 * r + 1], 255);
    HIPCHK(s, hipMemcpy(d.p, codes.data(), codes.size(), hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(d.rp, rp.data(), rp.size() * sizeof(float), hipMemcpyHostToDevice));
    return TW_OK;
  }
  if (c->coef16 == 3 && !ref_precision(c->precision)) {
    std::vector<uint16_t> codes((size_t)d.rows * d.pitch, 0);
    std::vector<float> rp((size_t)d.rows * 2);
    for (int r = 0; r < d.rows; r++)
      quantize_row(s->hcoef[k].data() + (size_t)r * d.cols, d.cols, codes.data() + (size_t)r * d.pitch, &rp[2 * r],
                   &rp[2 * r + 1], 65535);
    HIPCHK(s, hipMemcpy(d.p, codes.data(), codes.size() * 2, hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(d.rp, rp.data(), rp.size() * sizeof(float), hipMemcpyHostToDevice));
    return TW_OK;
  }
  const int cp = coef_store_precision(c);
  const size_t es = state_size(cp);
  std::vector<uint8_t> buf((size_t)d.rows * d.pitch * es, 0);
  for (int r = 0; r < d.rows; r++)
    to_storage(cp, s->hcoef[k].data() + (size_t)r * d.cols, buf.data() + (size_t)r * d.pitch * es, (size_t)d.cols);
  HIPCHK(s, hipMemcpy(d.p, buf.data(), buf.size(), hipMemcpyHostToDevice));
  return TW_OK;
}

// impedances Z = sqrt(M / binv) = rho c along characteristic sides
int upload_z(tw_solver* s) {
  const tw_config* c = &s->cfg;
  const bool el = c->physics == TW_ELASTIC;
  const int Mx = c->mx, My = c->my;
  for (int side = 0; side < 4; side++) {
    if (c->bc[side] != 'C') continue;
    const int axis = side / 2, right = side % 2;
    for (int f = TW_VX; f <= TW_VY; f++) {
      const bool normal = (axis == 0) == (f == TW_VX);
      if (!el && !normal) continue;
      const tw::DevField& F = s->plan.f[f];
      const int n = axis == 0 ? F.cols : F.rows;
      std::vector<double> z(n);
      const std::vector<double>& binv = s->hcoef[f == TW_VX ? TW_C_BINV_VX : TW_C_BINV_VY];
      for (int j = 0; j < n; j++) {
        const int a = axis == 0 ? (right ? F.rows - 1 : 0) : (right ? F.cols - 1 : 0);
        const size_t qv = axis == 0 ? (size_t)a * F.cols + j : (size_t)j * F.cols + a;
        double mod;
        if (normal) {
          const int am = axis == 0 ? (right ? Mx - 1 : 0) : (right ? My - 1 : 0);
          const size_t qm = axis == 0 ? (size_t)am * My + j : (size_t)j * My + am;
          mod = el ? s->hcoef[TW_C_LAMBDA][qm] + s->hcoef[TW_C_TWOMU][qm] : s->hcoef[TW_C_KAPPA][qm];
        } else {
          const int an = axis == 0 ? (right ? Mx : 0) : (right ? My : 0);
          const size_t qn = axis == 0 ? (size_t)an * (My + 1) + j : (size_t)j * (My + 1) + an;
          mod = s->hcoef[TW_C_MU][qn];
        }
        z[j] = sqrt(mod / binv[qv]);
      }
      void* dz = s->plan.z[f == TW_VX ? 0 : 1][side];
      if (c->precision == TW_FP64) {
        HIPCHK(s, hipMemcpy(dz, z.data(), n * 8, hipMemcpyHostToDevice));
      } else {
        std::vector<float> zf(n);
        for (int j = 0; j < n; j++) zf[j] = (float)z[j];
        HIPCHK(s, hipMemcpy(dz, zf.data(), n * 4, hipMemcpyHostToDevice));
      }
    }
  }
  return TW_OK;
}

int after_coef_change(tw_solver* s) {
  s->ep_valid = false;
  s->gen++;
  return upload_z(s);
}

void set_default_material(tw_solver* s) {
  const tw_config* c = &s->cfg;
  const double rho = 1., vp = c->physics == TW_ELASTIC ? 2. : 1., vs = 1.;
  for (int k = 0; k < 6; k++) {
    if (!coef_in(c->physics, k)) continue;
    double v;
    switch (k) {
      case TW_C_BINV_VX:
      case TW_C_BINV_VY: v = 1. / rho; break;
      case TW_C_MU: v = rho * vs * vs; break;
      case TW_C_TWOMU: v = 2 * rho * vs * vs; break;
      case TW_C_LAMBDA: v = rho * vp * vp - 2 * rho * vs * vs; break;
      default: v = rho * vp * vp; break;
    }
    std::fill(s->hcoef[k].begin(), s->hcoef[k].end(), round_coef(c, v));
  }
}

// the fused step's buffers: the written one becomes current
void swap_buffers(Plan& p) {
  for (int f = 0; f < tw::NF; f++) {
    std::swap(p.f[f].s, p.f[f].s2);
    std::swap(p.f[f].r, p.f[f].r2);
  }
  p.parity ^= 1;
  p.dstat = p.dstat_ip[p.parity];
}

tw::Pending stress_pending(tw_solver* s, int it_done, bool energy) {
  tw::Pending p;
  const tw_config* c = &s->cfg;
  if (energy) {
    p.part = s->plan.part_s;
    p.nparts = s->plan.nparts;
    p.out0 = s->plan.ep1;
    p.it0 = it_done;
    if (it_done + 1 < c->nt_max) {
      p.out1 = s->plan.ep0;
      p.it1 = it_done + 1;
    }
    p.scale = c->energy_form == 1 ? 1.0 : 0.5;
  }
  if (!s->quiet && it_done < c->nt_max && !s->rcv_s.empty()) {
    p.rcv = s->plan.rcv_s;
    p.nrcv = (int32_t)s->rcv_s.size();
    p.rcv_it = it_done;
  }
  return p;
}

}  // namespace

// ---------------------------------------------------------------- C-ABI
int tw::host_fail(int code, const char* msg) { return fail(nullptr, code, msg); }

extern "C" {

const char* tw_global_error(void) { return g_err.c_str(); }

int tw_get_config(const tw_solver* s, tw_config* cfg) {
  if (!s || !cfg) return fail(nullptr, TW_EINVAL, "null argument");
  *cfg = s->cfg;
  return TW_OK;
}
const char* tw_version(void) { return "twave 0.1 (gfx950)"; }

void tw_config_default(tw_config* c) {
  memset(c, 0, sizeof *c);
  c->physics = TW_ELASTIC;
  c->mx = c->my = 64;
  c->dx = 1.0 / 125;
  c->num_dx = 1;
  c->den_dx = 125;
  c->dt = 1e-4;
  memcpy(c->bc, "FFPP", 4);
  c->precision = TW_FP32;
  c->cpst = TW_CPST_0;
  c->record_energy = 1;
  c->nt_max = 1000;
  c->source_scaling = 1.0;
}

int tw_config_check(const tw_config* cfg) { return check_config(cfg, nullptr); }

int tw_field_shape(const tw_config* cfg, int32_t field, int32_t* rows, int32_t* cols) {
  if (!cfg || !rows || !cols) return fail(nullptr, TW_EINVAL, "null argument");
  if (!field_in(cfg->physics, field)) return fail(nullptr, TW_EINVAL, "field not present in this physics");
  shape_of(cfg, field_grid(field), rows, cols);
  return TW_OK;
}

int tw_locate(const tw_config* cfg, int32_t grid, int64_t num_x, int64_t den_x, int64_t num_y, int64_t den_y,
              int32_t* ix, int32_t* iy) {
  if (!cfg || !ix || !iy || grid < 0 || grid > 3) return fail(nullptr, TW_EINVAL, "bad argument");
  if (cfg->num_dx <= 0 || cfg->den_dx <= 0) return fail(nullptr, TW_EINVAL, "num_dx/den_dx not set");
  int32_t rx, ry;
  int e = locate_1d(num_x, den_x, cfg->num_dx, cfg->den_dx, grid_tx(grid), &rx);
  if (e) return fail(nullptr, e, "location is not a grid point along x");
  e = locate_1d(num_y, den_y, cfg->num_dx, cfg->den_dx, grid_ty(grid), &ry);
  if (e) return fail(nullptr, e, "location is not a grid point along y");
  int rows, cols;
  shape_of(cfg, grid, &rows, &cols);
  if (rx < 0 || rx >= rows || ry < 0 || ry >= cols) return fail(nullptr, TW_ERANGE, "location outside the grid");
  *ix = rx;
  *iy = ry;
  return TW_OK;
}

int tw_interpolate_model(const tw_config* cfg, int32_t grid, const double* model, int32_t nxm, int32_t nym,
                         int64_t num_dxm, int64_t den_dxm, double* out) {
  if (!cfg || !model || !out || grid < 0 || grid > 3) return fail(nullptr, TW_EINVAL, "bad argument");
  return interpolate(cfg, grid, model, nxm, nym, num_dxm, den_dxm, out);
}

int tw_interpolate_model_adjoint(const tw_config* cfg, int32_t grid, const double* values, int32_t nxm, int32_t nym,
                                 int64_t num_dxm, int64_t den_dxm, double* model_out) {
  if (!cfg || !values || !model_out || grid < 0 || grid > 3) return fail(nullptr, TW_EINVAL, "bad argument");
  return interpolate_adjoint(cfg, grid, values, nxm, nym, num_dxm, den_dxm, model_out);
}

double tw_ricker_increment(const tw_config* cfg, double f0, double t0, int32_t it, int32_t half) {
  return ricker_inc(cfg, f0, t0, it, half);
}

// grids of at most this many union points step with the per-point kernels
// unless TW_KERNEL says otherwise (about 1200^2)
constexpr int64_t kPointCells = 1500000;

int tw_create(const tw_config* cfg, tw_solver** out) {
  if (!out) return fail(nullptr, TW_EINVAL, "null output handle");
  *out = nullptr;
  int e = check_config(cfg, nullptr);
  if (e) return e;
  tw_solver* s = new tw_solver();
  s->cfg = *cfg;
  Plan& p = s->plan;
  const tw_config* c = &s->cfg;
  hipError_t he = hipSetDevice(c->device);
  if (he != hipSuccess) {
    hip_fail(s, he, "hipSetDevice");
    delete s;
    return TW_EDEVICE;
  }
  // set-up of the allocation-free entry point: its scratch address on this
  // device (looked up once per device, the device current)
  int32_t* scratch_ptr = nullptr;
  he = tw::scratch_address(c->device, &scratch_ptr);
  if (he != hipSuccess) {
    hip_fail(s, he, "hipGetSymbolAddress(scratch)");
    delete s;
    return TW_EDEVICE;
  }
  p.physics = c->physics;
  p.precision = c->precision;
  p.coef16 = c->coef16;
  p.cpst = c->cpst;
  p.mx = c->mx;
  p.my = c->my;
  p.perx = c->bc[0] == 'P';
  p.pery = c->bc[2] == 'P';
  memcpy(p.bc, c->bc, 4);
  p.energy_form = c->energy_form;
  p.dx = c->dx;
  if (const char* kv = getenv("TW_KERNEL")) {
    p.kver = atoi(kv) == 1 ? 1 : (atoi(kv) == 3 ? 3 : 2);
  } else if ((int64_t)(c->mx + 1) * (c->my + 1) <= kPointCells) {
    // on grids this small a step is latency bound: the per-point kernels (one
    // thread per point, all loads of a point in flight at once) beat waves
    // walking rows (512^2 elastic BF16 + K8: 0.031 against 0.074 ms per
    // step, 1024^2 0.068 against 0.080; at 2048^2 the streaming kernels run
    // 0.106 against 0.198)
    p.kver = 1;
    s->auto_point = true;
  }
  p.txc = tw::default_chunk_rows(c->precision, c->cpst);
  {
    // a wave walks the rows of its chunk one after another, so a grid with few
    // strips and chunks leaves the chip idle: shorter chunks (down to 4 rows)
    // until the core launch has a wave for every wave slot the core kernels
    // can hold (CUs x 4 SIMDs x their waves per SIMD: FP64 1, FP32 with a
    // residual array 3, acoustic 16-bit 8, otherwise 4)
    int cus = 256;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || cus <= 0)
      cus = 256;
    const bool sixteen = c->precision == TW_BF16 || c->precision == TW_FP16;
    const int wpe = c->precision == TW_FP64 ? 1
                    : (sixteen && c->physics == TW_ACOUSTIC) ? 8
                    : (c->precision == TW_FP32 && c->cpst != TW_CPST_0) ? 3 : 4;
    const int64_t slots = (int64_t)cus * 4 * wpe;
    const int64_t strips = ((int64_t)c->my + 1 + 255) / 256;
    const int64_t want = ((int64_t)c->mx * strips + slots - 1) / slots;
    if (want < p.txc) p.txc = (int32_t)std::max<int64_t>(4, want);
  }
  if (const char* tx = getenv("TW_TXC"))
    if (atoi(tx) > 0) p.txc = atoi(tx);
  if (const char* mb = getenv("TW_MERGE_BANDS")) p.merge_bands = atoi(mb) != 0;
  int pad = 0;
  if (const char* pp = getenv("TW_PITCH_PAD")) pad = atoi(pp) > 0 ? (atoi(pp) + 3) / 4 * 4 : 0;
  build_tables(c, &p.t);
  const size_t es = state_size(c->precision);
  auto bail = [&](int code) {
    tw_destroy(s);
    return code;
  };
  for (int f = 0; f < 6; f++) {
    if (!field_in(c->physics, f)) continue;
    tw::DevField& d = p.f[f];
    shape_of(c, field_grid(f), &d.rows, &d.cols);
    d.pitch = tw::pitch_of(d.cols, pad);
    const size_t bytes = (size_t)d.rows * d.pitch * es;
    const size_t rbytes = (size_t)d.rows * d.pitch * array_size(c, 1);
    if ((e = dalloc(s, &d.s, bytes))) return bail(e);
    if (rmode(c->cpst) && (e = dalloc(s, &d.r, rbytes))) return bail(e);
    if (cmode(c->cpst) && (e = dalloc(s, &d.c, bytes))) return bail(e);
  }
  p.fused = p.kver == 3 && tw::fused_ok(p);
  if (p.fused)
    for (int f = 0; f < 6; f++) {  // the other buffer of each ping-pong pair
      if (!field_in(c->physics, f)) continue;
      tw::DevField& d = p.f[f];
      const size_t bytes = (size_t)d.rows * d.pitch * es;
      if ((e = dalloc(s, &d.s2, bytes))) return bail(e);
      if (rmode(c->cpst) && (e = dalloc(s, &d.r2, (size_t)d.rows * d.pitch * array_size(c, 1)))) return bail(e);
    }
  const size_t ks = coef_size(c->precision, c->coef16);
  for (int k = 0; k < 6; k++) {
    if (!coef_in(c->physics, k)) continue;
    tw::DevCoef& d = p.k[k];
    shape_of(c, coef_grid(k), &d.rows, &d.cols);
    d.pitch = tw::pitch_of(d.cols, pad);
    if ((e = dalloc(s, &d.p, (size_t)d.rows * d.pitch * ks))) return bail(e);
    if (c->coef16 >= 2 && (e = dalloc(s, &d.rp, (size_t)d.rows * 2 * sizeof(float)))) return bail(e);
    s->hcoef[k].assign((size_t)d.rows * d.cols, 0.0);
  }
  // characteristic SAT buffers
  const size_t as = arith_size(c->precision);
  for (int side = 0; side < 4; side++) {
    if (c->bc[side] != 'C') continue;
    const int axis = side / 2;
    for (int f = TW_VX; f <= TW_VY; f++) {
      const bool normal = (axis == 0) == (f == TW_VX);
      if (c->physics == TW_ACOUSTIC && !normal) continue;
      const tw::DevField& F = p.f[f];
      const int n = axis == 0 ? F.cols : F.rows;
      if ((e = dalloc(s, &p.z[f == TW_VX ? 0 : 1][side], (size_t)n * as))) return bail(e);
      if (!normal) {
        if ((e = dalloc(s, &p.uhat[f == TW_VX ? 0 : 1][side], (size_t)n * as))) return bail(e);
        s->need_uhat = true;
      }
    }
  }
  // records and partial sums
  p.nt_max = c->nt_max;
  const size_t nt = (size_t)(c->nt_max > 0 ? c->nt_max : 1);
  if ((e = dalloc(s, (void**)&p.ek, nt * 8)) || (e = dalloc(s, (void**)&p.ep0, nt * 8)) ||
      (e = dalloc(s, (void**)&p.ep1, nt * 8)))
    return bail(e);
  p.nparts = p.fused ? tw::fused_parts(p) : tw::half_parts(p);
  const size_t npart_max = (size_t)std::max(p.nparts, tw::energy_parts(p));
  s->part_cap = npart_max;
  if ((e = dalloc(s, (void**)&p.part_v, npart_max * 8)) || (e = dalloc(s, (void**)&p.part_s, npart_max * 8)))
    return bail(e);
  // device copies of the static kernel arguments: in place, and (fused step)
  // read-current / write-other, for each parity of the ping-pong pairs
  for (int par = 0; par < (p.fused ? 2 : 1); par++) {
    for (int pp = 0; pp < (p.fused ? 2 : 1); pp++) {
      const size_t sb = tw::stat_fill(p, nullptr, pp);
      std::vector<uint8_t> hb(sb);
      tw::stat_fill(p, hb.data(), pp);
      void** dst = pp ? &p.dstat_pp[par] : &p.dstat_ip[par];
      if ((e = dalloc(s, dst, sb))) return bail(e);
      he = hipMemcpy(*dst, hb.data(), sb, hipMemcpyHostToDevice);
      if (he != hipSuccess) {
        hip_fail(s, he, "hipMemcpy");
        return bail(TW_EDEVICE);
      }
    }
    if (p.fused) swap_buffers(p);
  }
  p.dstat = p.dstat_ip[p.parity];
  he = hipStreamCreate(&s->stream);
  if (he != hipSuccess) {
    hip_fail(s, he, "hipStreamCreate");
    return bail(TW_EDEVICE);
  }
  s->own_stream = true;
  set_default_material(s);
  for (int k = 0; k < 6; k++)
    if (coef_in(c->physics, k) && (e = upload_coef(s, k))) return bail(e);
  if ((e = after_coef_change(s))) return bail(e);
  *out = s;
  return TW_OK;
}

int tw_destroy(tw_solver* s) {
  if (!s) return TW_OK;
  (void)hipSetDevice(s->cfg.device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (hipEvent_t e : s->prof_ev) (void)hipEventDestroy(e);
  for (void* p : s->allocs) (void)hipFree(p);
  if (s->own_stream && s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return TW_OK;
}

const char* tw_last_error(const tw_solver* s) { return s ? s->err.c_str() : g_err.c_str(); }

int tw_set_model(tw_solver* s, const double* rho, const double* vp, const double* vs, int32_t nxm, int32_t nym,
                 int64_t num_dxm, int64_t den_dxm) {
  if (!s || !rho || !vp || (s->cfg.physics == TW_ELASTIC && !vs)) return fail(s ? &s->err : nullptr, TW_EINVAL, "null argument");
  (void)hipSetDevice(s->cfg.device);
  const tw_config* c = &s->cfg;
  if (nxm < 1 || nym < 1) return fail(&s->err, TW_EINVAL, "model grid must have at least 2x2 nodes");
  const size_t nm = (size_t)(nxm + 1) * (nym + 1);
  for (size_t i = 0; i < nm; i++)
    if (!(rho[i] > 0) || !(vp[i] > 0) || !std::isfinite(rho[i]) || !std::isfinite(vp[i]) ||
        (c->physics == TW_ELASTIC && (!(vs[i] >= 0) || !std::isfinite(vs[i]))))
      return fail(&s->err, TW_EINVAL, "model values must be finite, rho and vp positive");
  std::vector<double> a(nm), b(nm);
  for (int k = 0; k < 6; k++) {
    if (!coef_in(c->physics, k)) continue;
    const int g = coef_grid(k);
    std::vector<double> out(s->hcoef[k].size());
    int e;
    switch (k) {
      case TW_C_BINV_VX:
      case TW_C_BINV_VY:  // make_density_reciprocal: 1 / (interpolated rho)
        e = interpolate(c, g, rho, nxm, nym, num_dxm, den_dxm, out.data());
        for (double& v : out) v = 1. / v;
        break;
      case TW_C_MU:
        for (size_t i = 0; i < nm; i++) a[i] = rho[i] * vs[i] * vs[i];
        e = interpolate(c, g, a.data(), nxm, nym, num_dxm, den_dxm, out.data());
        break;
      case TW_C_TWOMU:
        for (size_t i = 0; i < nm; i++) a[i] = 2 * rho[i] * vs[i] * vs[i];
        e = interpolate(c, g, a.data(), nxm, nym, num_dxm, den_dxm, out.data());
        break;
      case TW_C_LAMBDA:  // lambda = rho vp^2 - 2 mu on the model nodes
        for (size_t i = 0; i < nm; i++) {
          b[i] = 2 * rho[i] * vs[i] * vs[i];
          a[i] = rho[i] * vp[i] * vp[i] - b[i];
        }
        e = interpolate(c, g, a.data(), nxm, nym, num_dxm, den_dxm, out.data());
        break;
      default:  // kappa = rho vp^2
        for (size_t i = 0; i < nm; i++) a[i] = rho[i] * vp[i] * vp[i];
        e = interpolate(c, g, a.data(), nxm, nym, num_dxm, den_dxm, out.data());
        break;
    }
    if (e) {
      s->err = g_err;
      return e;
    }
    for (size_t i = 0; i < out.size(); i++) s->hcoef[k][i] = round_coef(c, out[i]);
    if ((e = upload_coef(s, k))) return e;
  }
  return after_coef_change(s);
}

int tw_set_coef(tw_solver* s, int32_t k, const double* v) {
  if (!s || !v) return fail(s ? &s->err : nullptr, TW_EINVAL, "null argument");
  if (!coef_in(s->cfg.physics, k)) return fail(&s->err, TW_EINVAL, "coefficient not used by this physics");
  (void)hipSetDevice(s->cfg.device);
  const size_t n = s->hcoef[k].size();
  // validate the whole array first: a refused call leaves the coefficients as they were
  for (size_t i = 0; i < n; i++) {
    if (!std::isfinite(v[i])) return fail(&s->err, TW_EINVAL, "coefficient must be finite");
    if (!(v[i] > 0) && k != TW_C_LAMBDA) return fail(&s->err, TW_EINVAL, "coefficient must be positive");
  }
  for (size_t i = 0; i < n; i++) s->hcoef[k][i] = round_coef(&s->cfg, v[i]);
  int e = upload_coef(s, k);
  if (e) return e;
  return after_coef_change(s);
}

int tw_get_coef(tw_solver* s, int32_t k, double* v) {
  if (!s || !v) return fail(s ? &s->err : nullptr, TW_EINVAL, "null argument");
  if (!coef_in(s->cfg.physics, k)) return fail(&s->err, TW_EINVAL, "coefficient not used by this physics");
  memcpy(v, s->hcoef[k].data(), s->hcoef[k].size() * 8);
  return TW_OK;
}

static int check_source_point(tw_solver* s, int32_t grid, int32_t ix, int32_t iy) {
  const tw_config* c = &s->cfg;
  if (grid < 0 || grid > 3) return fail(&s->err, TW_EINVAL, "grid id");
  if (c->physics == TW_ACOUSTIC && grid == TW_GRID_NN) return fail(&s->err, TW_EINVAL, "no NN grid in acoustic");
  int rows, cols;
  shape_of(c, grid, &rows, &cols);
  if (ix < 0 || ix >= rows || iy < 0 || iy >= cols) return fail(&s->err, TW_ERANGE, "source outside the grid");
  // forward.cpp:95-98: no source on a weak (SBP) node boundary
  if (c->bc[0] != 'P' && grid_tx(grid) && (ix == 0 || ix == rows - 1))
    return fail(&s->err, TW_ERANGE, "source on a weak boundary along x");
  if (c->bc[2] != 'P' && grid_ty(grid) && (iy == 0 || iy == cols - 1))
    return fail(&s->err, TW_ERANGE, "source on a weak boundary along y");
  return TW_OK;
}

int tw_set_source_ricker(tw_solver* s, int32_t grid, int32_t ix, int32_t iy, double f0, double t0) {
  if (!s) return fail(nullptr, TW_EINVAL, "null handle");
  int e = check_source_point(s, grid, ix, iy);
  if (e) return e;
  const int half = !(grid == TW_GRID_MN || grid == TW_GRID_NM);
  s->src_inc.resize((size_t)s->cfg.nt_max);
  for (int it = 0; it < s->cfg.nt_max; it++) s->src_inc[it] = ricker_inc(&s->cfg, f0, t0, it, half);
  s->src_grid = grid;
  s->src_ix = ix;
  s->src_iy = iy;
  s->gen++;
  return TW_OK;
}

int tw_set_source_wavelet(tw_solver* s, int32_t grid, int32_t ix, int32_t iy, const double* inc, int32_t n) {
  if (!s || (!inc && n > 0) || n < 0) return fail(s ? &s->err : nullptr, TW_EINVAL, "bad argument");
  int e = check_source_point(s, grid, ix, iy);
  if (e) return e;
  s->src_inc.assign(inc, inc + n);
  s->src_grid = grid;
  s->src_ix = ix;
  s->src_iy = iy;
  s->gen++;
  return TW_OK;
}

int tw_clear_source(tw_solver* s) {
  if (!s) return fail(nullptr, TW_EINVAL, "null handle");
  s->src_grid = -1;
  s->src_inc.clear();
  s->gen++;
  return TW_OK;
}

int tw_set_receivers(tw_solver* s, int32_t n, const int32_t* field, const int32_t* ix, const int32_t* iy) {
  if (!s || n < 0 || (n > 0 && (!field || !ix || !iy))) return fail(s ? &s->err : nullptr, TW_EINVAL, "bad argument");
  (void)hipSetDevice(s->cfg.device);
  std::vector<tw::Receiver> v, st;
  for (int r = 0; r < n; r++) {
    if (!field_in(s->cfg.physics, field[r])) return fail(&s->err, TW_EINVAL, "receiver field not in this physics");
    int rows, cols;
    shape_of(&s->cfg, field_grid(field[r]), &rows, &cols);
    if (ix[r] < 0 || ix[r] >= rows || iy[r] < 0 || iy[r] >= cols) return fail(&s->err, TW_ERANGE, "receiver outside the grid");
    tw::Receiver rc{field[r], ix[r], iy[r], r};
    (is_velocity(field[r]) ? v : st).push_back(rc);
  }
  HIPCHK(s, hipStreamSynchronize(s->stream));
  // records [n][nt_max]
  dfree(s, (void**)&s->plan.rec);
  dfree(s, (void**)&s->plan.rcv_v);
  dfree(s, (void**)&s->plan.rcv_s);
  int e;
  const size_t recn = (size_t)(n > 0 ? n : 1) * (size_t)(s->cfg.nt_max > 0 ? s->cfg.nt_max : 1);
  if ((e = dalloc(s, (void**)&s->plan.rec, recn * 8))) return e;
  if ((e = dalloc(s, (void**)&s->plan.rcv_v, (v.size() + 1) * sizeof(tw::Receiver)))) return e;
  if ((e = dalloc(s, (void**)&s->plan.rcv_s, (st.size() + 1) * sizeof(tw::Receiver)))) return e;
  if (!v.empty()) HIPCHK(s, hipMemcpy(s->plan.rcv_v, v.data(), v.size() * sizeof(tw::Receiver), hipMemcpyHostToDevice));
  if (!st.empty()) HIPCHK(s, hipMemcpy(s->plan.rcv_s, st.data(), st.size() * sizeof(tw::Receiver), hipMemcpyHostToDevice));
  s->rcv_v = v;
  s->rcv_s = st;
  s->plan.nrcv_v = (int32_t)v.size();
  s->plan.nrcv_s = (int32_t)st.size();
  s->nrcv = n;
  s->gen++;
  return TW_OK;
}

int tw_get_receivers(tw_solver* s, int32_t cap, int32_t* n, int32_t* field, int32_t* ix, int32_t* iy) {
  if (!s || !n || cap < 0) return fail(s ? &s->err : nullptr, TW_EINVAL, "bad argument");
  *n = s->nrcv;
  if (!field && !ix && !iy) return TW_OK;
  if (cap < s->nrcv) return fail(&s->err, TW_ERANGE, "more receivers than cap");
  for (const auto* list : {&s->rcv_v, &s->rcv_s})
    for (const tw::Receiver& r : *list) {
      if (field) field[r.row] = r.field;
      if (ix) ix[r.row] = r.ix;
      if (iy) iy[r.row] = r.iy;
    }
  return TW_OK;
}

int tw_step(tw_solver* s, int32_t nsteps) {
  if (!s || nsteps < 0) return fail(s ? &s->err : nullptr, TW_EINVAL, "bad argument");
  (void)hipSetDevice(s->cfg.device);
  const tw_config* c = &s->cfg;
  Plan& p = s->plan;
  const bool vsrc = s->src_grid == TW_GRID_MN || s->src_grid == TW_GRID_NM;
  tw::Pending carry;  // stress work of the previous step, done by the next launch
  bool have_carry = false;
  if (p.fused) {
    for (int k = 0; k < nsteps; k++) {
      const int it = s->it;
      const bool rec = !s->quiet && c->record_energy && it < c->nt_max;
      if (rec && !s->ep_valid) {
        HIPCHK(s, tw::launch_stress_energy(p, s->stream));
        carry = tw::Pending();
        carry.part = p.part_s;
        carry.nparts = tw::energy_parts(p);
        carry.out0 = p.ep0;
        carry.it0 = it;
        carry.scale = c->energy_form == 1 ? 1.0 : 0.5;
        have_carry = true;
        s->ep_valid = true;
      }
      if (s->need_uhat) HIPCHK(s, tw::launch_uhat(p, s->stream));
      const bool src = s->src_grid >= 0 && it < (int)s->src_inc.size();
      auto ctl = [&](bool on) {
        tw::StepCtl x;
        x.it = it;
        x.energy = rec;
        if (on) {
          x.src_on = 1;
          x.src_grid = s->src_grid;
          x.src_ix = s->src_ix;
          x.src_iy = s->src_iy;
          x.src_inc = s->src_inc[it];
        }
        return x;
      };
      tw::StepCtl cv = ctl(src && vsrc);
      if (have_carry) cv.pend = carry;
      HIPCHK(s, tw::launch_fused(p, 0, cv, s->stream));
      const bool prof = s->prof_n < s->prof_max;
      if (prof) HIPCHK(s, hipEventRecord(s->prof_ev[4 * s->prof_n], s->stream));
      HIPCHK(s, tw::launch_fused(p, 1, ctl(src), s->stream));
      if (prof) HIPCHK(s, hipEventRecord(s->prof_ev[4 * s->prof_n + 1], s->stream));
      if (prof) s->prof_n++;
      tw::StepCtl cs = ctl(src && !vsrc);
      if (rec) {
        cs.pend.part = p.part_v;
        cs.pend.nparts = p.nparts;
        cs.pend.out0 = p.ek;
        cs.pend.it0 = it;
        cs.pend.scale = c->energy_form == 1 ? 1.0 : 0.5;
      }
      if (!s->quiet && it < c->nt_max && !s->rcv_v.empty()) {
        cs.pend.rcv = p.rcv_v;
        cs.pend.nrcv = (int32_t)s->rcv_v.size();
        cs.pend.rcv_it = it;
      }
      HIPCHK(s, tw::launch_fused(p, 2, cs, s->stream));
      carry = stress_pending(s, it, rec);
      have_carry = true;
      swap_buffers(p);
      s->it++;
    }
    if (have_carry && (carry.nparts > 0 || carry.nrcv > 0)) HIPCHK(s, tw::launch_finalize(p, carry, s->stream));
    return TW_OK;
  }
  for (int k = 0; k < nsteps; k++) {
    const int it = s->it;
    const bool rec = !s->quiet && c->record_energy && it < c->nt_max;
    if (rec && !s->ep_valid) {
      HIPCHK(s, tw::launch_stress_energy(p, s->stream));
      carry = tw::Pending();
      carry.part = p.part_s;
      carry.nparts = tw::energy_parts(p);
      carry.out0 = p.ep0;
      carry.it0 = it;
      carry.scale = c->energy_form == 1 ? 1.0 : 0.5;
      have_carry = true;
      s->ep_valid = true;
    }
    if (s->need_uhat) HIPCHK(s, tw::launch_uhat(p, s->stream));
    tw::StepCtl cv;
    cv.it = it;
    cv.energy = rec;
    if (s->src_grid >= 0 && vsrc && it < (int)s->src_inc.size()) {
      cv.src_on = 1;
      cv.src_grid = s->src_grid;
      cv.src_ix = s->src_ix;
      cv.src_iy = s->src_iy;
      cv.src_inc = s->src_inc[it];
    }
    if (have_carry) cv.pend = carry;
    const bool prof = s->prof_n < s->prof_max;
    for (int i = 0; i < 4; i++) p.prof[i] = prof ? s->prof_ev[4 * s->prof_n + i] : nullptr;
    HIPCHK(s, tw::launch_half(p, 0, cv, s->stream));
    tw::StepCtl cs;
    cs.it = it;
    cs.energy = rec;
    if (s->src_grid >= 0 && !vsrc && it < (int)s->src_inc.size()) {
      cs.src_on = 1;
      cs.src_grid = s->src_grid;
      cs.src_ix = s->src_ix;
      cs.src_iy = s->src_iy;
      cs.src_inc = s->src_inc[it];
    }
    if (rec) {
      cs.pend.part = p.part_v;
      cs.pend.nparts = p.nparts;
      cs.pend.out0 = p.ek;
      cs.pend.it0 = it;
      cs.pend.scale = c->energy_form == 1 ? 1.0 : 0.5;
    }
    if (!s->quiet && it < c->nt_max && !s->rcv_v.empty()) {
      cs.pend.rcv = p.rcv_v;
      cs.pend.nrcv = (int32_t)s->rcv_v.size();
      cs.pend.rcv_it = it;
    }
    HIPCHK(s, tw::launch_half(p, 1, cs, s->stream));
    for (int i = 0; i < 4; i++) p.prof[i] = nullptr;
    if (prof) s->prof_n++;
    carry = stress_pending(s, it, rec);
    have_carry = true;
    s->it++;
  }
  if (have_carry && (carry.nparts > 0 || carry.nrcv > 0)) HIPCHK(s, tw::launch_finalize(p, carry, s->stream));
  return TW_OK;
}

static void prof_release(tw_solver* s) {
  for (hipEvent_t e : s->prof_ev) (void)hipEventDestroy(e);
  s->prof_ev.clear();
  s->prof_n = s->prof_max = 0;
}

int tw_profile_begin(tw_solver* s, int32_t max_steps) {
  if (!s || max_steps < 1 || max_steps > (1 << 20)) return fail(s ? &s->err : nullptr, TW_EINVAL, "bad argument");
  (void)hipSetDevice(s->cfg.device);
  HIPCHK(s, hipStreamSynchronize(s->stream));
  prof_release(s);
  s->prof_ev.resize((size_t)4 * max_steps, nullptr);
  for (auto& e : s->prof_ev) {
    hipError_t r = hipEventCreate(&e);
    if (r != hipSuccess) {
      prof_release(s);
      return hip_fail(s, r, "hipEventCreate");
    }
  }
  s->prof_max = max_steps;
  return TW_OK;
}

int tw_profile_end(tw_solver* s, double* ms_velocity, double* ms_stress, int32_t* nsteps) {
  if (!s || !ms_velocity || !ms_stress || !nsteps) return fail(s ? &s->err : nullptr, TW_EINVAL, "null argument");
  (void)hipSetDevice(s->cfg.device);
  HIPCHK(s, hipStreamSynchronize(s->stream));
  double v = 0.0, st = 0.0;
  const bool fused = s->plan.fused;
  for (int k = 0; k < s->prof_n; k++) {
    float a = 0.f, b = 0.f;
    HIPCHK(s, hipEventElapsedTime(&a, s->prof_ev[4 * k], s->prof_ev[4 * k + 1]));
    if (!fused) HIPCHK(s, hipEventElapsedTime(&b, s->prof_ev[4 * k + 2], s->prof_ev[4 * k + 3]));
    v += a;
    st += b;
  }
  *ms_velocity = v;
  *ms_stress = st;
  *nsteps = s->prof_n;
  prof_release(s);
  return TW_OK;
}

int tw_data_misfit(tw_solver* s, const double* data, int32_t nt, double* per_record, double* total) {
  if (!s || (!data && s->nrcv > 0) || !total || nt < 0) return fail(s ? &s->err : nullptr, TW_EINVAL, "bad argument");
  const int have = s->it < s->cfg.nt_max ? s->it : s->cfg.nt_max;
  if (nt > have) return fail(&s->err, TW_ERANGE, "more data steps than recorded steps");
  std::vector<double> rec((size_t)(s->nrcv > 0 ? s->nrcv : 1) * (have > 0 ? have : 1));
  int32_t got = 0;
  int e = tw_read_receivers(s, rec.data(), &got);
  if (e) return e;
  double sum = 0.0;
  for (int r = 0; r < s->nrcv; r++) {
    double m = 0.0;
    for (int t = 0; t < nt; t++) {
      const double d = rec[(size_t)r * got + t] - data[(size_t)r * nt + t];
      m += d * d;
    }
    if (per_record) per_record[r] = m;
    sum += m;
  }
  *total = sum;
  return TW_OK;
}

int tw_synchronize(tw_solver* s) {
  if (!s) return fail(nullptr, TW_EINVAL, "null handle");
  (void)hipSetDevice(s->cfg.device);
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return TW_OK;
}

int tw_current_step(tw_solver* s, int32_t* it) {
  if (!s || !it) return fail(s ? &s->err : nullptr, TW_EINVAL, "null argument");
  *it = s->it;
  return TW_OK;
}

int tw_read_receivers(tw_solver* s, double* out, int32_t* nt) {
  if (!s) return fail(nullptr, TW_EINVAL, "null handle");
  (void)hipSetDevice(s->cfg.device);
  const int n = s->it < s->cfg.nt_max ? s->it : s->cfg.nt_max;
  if (nt) *nt = n;
  if (!out || s->nrcv == 0) return TW_OK;
  HIPCHK(s, hipStreamSynchronize(s->stream));
  std::vector<double> buf((size_t)s->nrcv * s->cfg.nt_max);
  HIPCHK(s, hipMemcpy(buf.data(), s->plan.rec, buf.size() * 8, hipMemcpyDeviceToHost));
  for (int r = 0; r < s->nrcv; r++)
    for (int t = 0; t < n; t++) out[(size_t)r * n + t] = s->cfg.source_scaling * buf[(size_t)r * s->cfg.nt_max + t];
  return TW_OK;
}

int tw_read_energy(tw_solver* s, double* ek, double* ep0, double* ep1, int32_t* nt) {
  if (!s) return fail(nullptr, TW_EINVAL, "null handle");
  (void)hipSetDevice(s->cfg.device);
  const int n = s->it < s->cfg.nt_max ? s->it : s->cfg.nt_max;
  if (nt) *nt = n;
  if (n == 0) return TW_OK;
  HIPCHK(s, hipStreamSynchronize(s->stream));
  const double s2 = s->cfg.source_scaling * s->cfg.source_scaling;
  std::vector<double> b(n);
  double* outs[3] = {ek, ep0, ep1};
  double* devs[3] = {s->plan.ek, s->plan.ep0, s->plan.ep1};
  for (int i = 0; i < 3; i++) {
    if (!outs[i]) continue;
    HIPCHK(s, hipMemcpy(b.data(), devs[i], (size_t)n * 8, hipMemcpyDeviceToHost));
    for (int t = 0; t < n; t++) outs[i][t] = s2 * b[t];
  }
  return TW_OK;
}

static int copy_field(tw_solver* s, int32_t field, int which, double* out, const double* in) {
  if (!s || (!out && !in)) return fail(s ? &s->err : nullptr, TW_EINVAL, "null argument");
  if (!field_in(s->cfg.physics, field)) return fail(&s->err, TW_EINVAL, "field not present in this physics");
  (void)hipSetDevice(s->cfg.device);
  const tw::DevField& d = s->plan.f[field];
  void* dev = which == 0 ? d.s : (which == 1 ? d.r : d.c);
  const size_t es = array_size(&s->cfg, which);
  HIPCHK(s, hipStreamSynchronize(s->stream));
  if (!dev) {
    if (out) memset(out, 0, (size_t)d.rows * d.cols * 8);
    return in ? fail(&s->err, TW_ESTATE, "no companion array in this compensation mode") : TW_OK;
  }
  std::vector<uint8_t> buf((size_t)d.rows * d.cols * es);
  if (out) {
    HIPCHK(s, hipMemcpy2D(buf.data(), d.cols * es, dev, d.pitch * es, d.cols * es, d.rows, hipMemcpyDeviceToHost));
    if (es == 1) {  // K8 residual codes
      for (size_t i = 0; i < buf.size(); i++) out[i] = (double)(int8_t)buf[i];
    } else {
      from_storage(s->cfg.precision, buf.data(), out, (size_t)d.rows * d.cols);
    }
  } else {
    if (es == 1) return fail(&s->err, TW_EINVAL, "K8 residual codes are not set from values");
    to_storage(s->cfg.precision, in, buf.data(), (size_t)d.rows * d.cols);
    HIPCHK(s, hipMemcpy2D(dev, d.pitch * es, buf.data(), d.cols * es, d.cols * es, d.rows, hipMemcpyHostToDevice));
    s->ep_valid = false;
  }
  return TW_OK;
}

int tw_get_field(tw_solver* s, int32_t field, double* out) { return copy_field(s, field, 0, out, nullptr); }
int tw_set_field(tw_solver* s, int32_t field, const double* in) { return copy_field(s, field, 0, nullptr, in); }
int tw_get_companion(tw_solver* s, int32_t field, double* out) {
  if (!s) return fail(nullptr, TW_EINVAL, "null handle");
  return copy_field(s, field, cmode(s->cfg.cpst) ? 2 : 1, out, nullptr);
}

int tw_get_field_compensated(tw_solver* s, int32_t field, double* out) {
  if (!s || !out) return fail(s ? &s->err : nullptr, TW_EINVAL, "null argument");
  int e = tw_get_field(s, field, out);
  if (e || s->cfg.cpst == TW_CPST_0) return e;
  int rows, cols;
  shape_of(&s->cfg, field_grid(field), &rows, &cols);
  std::vector<double> comp((size_t)rows * cols);
  if ((e = tw_get_companion(s, field, comp.data()))) return e;
  if (k8mode(s->cfg.cpst)) {
    // code q -> q * 2^(max(e, 1) - bias), e the stored value's exponent field
    // (BF16: bias 142, exponent bias 127; FP16: bias 33, exponent bias 15)
    const int bias = s->cfg.precision == TW_BF16 ? 142 : 33, ebias = s->cfg.precision == TW_BF16 ? 127 : 15;
    for (size_t i = 0; i < comp.size(); i++) {
      int ex = 0;
      if (out[i] != 0.0 && std::isfinite(out[i])) {
        (void)std::frexp(out[i], &ex);
        ex = ex - 1 + ebias;  // biased exponent of a normal value
      }
      out[i] += std::ldexp(comp[i], std::max(ex, 1) - bias);
    }
  } else {
    for (size_t i = 0; i < comp.size(); i++) out[i] += comp[i];
  }
  return TW_OK;
}

int tw_reset(tw_solver* s) {
  if (!s) return fail(nullptr, TW_EINVAL, "null handle");
  (void)hipSetDevice(s->cfg.device);
  HIPCHK(s, hipStreamSynchronize(s->stream));
  for (int f = 0; f < 6; f++) {
    const tw::DevField& d = s->plan.f[f];
    void* arr[3] = {d.s, d.r, d.c};
    for (int w = 0; w < 3; w++)
      if (arr[w]) HIPCHK(s, hipMemset(arr[w], 0, (size_t)d.rows * d.pitch * array_size(&s->cfg, w)));
  }
  const size_t nt = (size_t)(s->cfg.nt_max > 0 ? s->cfg.nt_max : 1);
  for (double* p : {s->plan.ek, s->plan.ep0, s->plan.ep1}) HIPCHK(s, hipMemset(p, 0, nt * 8));
  if (s->plan.rec) HIPCHK(s, hipMemset(s->plan.rec, 0, (size_t)(s->nrcv > 0 ? s->nrcv : 1) * nt * 8));
  s->it = 0;
  s->ep_valid = false;
  return TW_OK;
}

// checkpoint layout: magic, version, step, then for each present array of
// each field (s, r, c) its rows*cols storage elements.
static const uint64_t kMagic = 0x3158414d56415754ull;  // "TWAVMAX1"

int tw_checkpoint_size(tw_solver* s, size_t* bytes) {
  if (!s || !bytes) return fail(s ? &s->err : nullptr, TW_EINVAL, "null argument");
  size_t n = 3 * sizeof(uint64_t);
  for (int f = 0; f < 6; f++) {
    const tw::DevField& d = s->plan.f[f];
    void* arr[3] = {d.s, d.r, d.c};
    for (int w = 0; w < 3; w++)
      if (arr[w]) n += (size_t)d.rows * d.cols * array_size(&s->cfg, w);
  }
  *bytes = n;
  return TW_OK;
}

int tw_checkpoint_save(tw_solver* s, void* buf, size_t bytes) {
  size_t need;
  int e = tw_checkpoint_size(s, &need);
  if (e) return e;
  if (!buf || bytes < need) return fail(&s->err, TW_EINVAL, "checkpoint buffer too small");
  (void)hipSetDevice(s->cfg.device);
  HIPCHK(s, hipStreamSynchronize(s->stream));
  uint64_t* h = (uint64_t*)buf;
  h[0] = kMagic;
  h[1] = ((uint64_t)s->cfg.precision << 32) | (uint64_t)s->cfg.cpst;
  h[2] = (uint64_t)s->it;
  uint8_t* o = (uint8_t*)buf + 3 * sizeof(uint64_t);
  for (int f = 0; f < 6; f++) {
    const tw::DevField& d = s->plan.f[f];
    void* arr[3] = {d.s, d.r, d.c};
    for (int w = 0; w < 3; w++) {
      void* p = arr[w];
      if (!p) continue;
      const size_t es = array_size(&s->cfg, w);
      HIPCHK(s, hipMemcpy2D(o, d.cols * es, p, d.pitch * es, d.cols * es, d.rows, hipMemcpyDeviceToHost));
      o += (size_t)d.rows * d.cols * es;
    }
  }
  return TW_OK;
}

int tw_checkpoint_load(tw_solver* s, const void* buf, size_t bytes) {
  size_t need;
  int e = tw_checkpoint_size(s, &need);
  if (e) return e;
  if (!buf || bytes < need) return fail(&s->err, TW_EINVAL, "checkpoint buffer too small");
  const uint64_t* h = (const uint64_t*)buf;
  if (h[0] != kMagic || h[1] != (((uint64_t)s->cfg.precision << 32) | (uint64_t)s->cfg.cpst))
    return fail(&s->err, TW_EINVAL, "checkpoint does not match this solver");
  (void)hipSetDevice(s->cfg.device);
  HIPCHK(s, hipStreamSynchronize(s->stream));
  const uint8_t* o = (const uint8_t*)buf + 3 * sizeof(uint64_t);
  for (int f = 0; f < 6; f++) {
    const tw::DevField& d = s->plan.f[f];
    void* arr[3] = {d.s, d.r, d.c};
    for (int w = 0; w < 3; w++) {
      void* p = arr[w];
      if (!p) continue;
      const size_t es = array_size(&s->cfg, w);
      HIPCHK(s, hipMemcpy2D(p, d.pitch * es, o, d.cols * es, d.cols * es, d.rows, hipMemcpyHostToDevice));
      o += (size_t)d.rows * d.cols * es;
    }
  }
  s->it = (int)h[2];
  s->ep_valid = false;
  return TW_OK;
}

int tw_set_stream(tw_solver* s, void* stream) {
  if (!s) return fail(nullptr, TW_EINVAL, "null handle");
  (void)hipSetDevice(s->cfg.device);
  HIPCHK(s, hipStreamSynchronize(s->stream));
  if (s->own_stream) (void)hipStreamDestroy(s->stream);
  s->stream = (hipStream_t)stream;
  s->own_stream = false;
  return TW_OK;
}

int tw_check_finite(tw_solver* s) {
  if (!s) return fail(nullptr, TW_EINVAL, "null handle");
  (void)hipSetDevice(s->cfg.device);
  int32_t* flag = nullptr;
  HIPCHK(s, hipMalloc((void**)&flag, sizeof(int32_t)));
  HIPCHK(s, hipMemsetAsync(flag, 0, sizeof(int32_t), s->stream));
  hipError_t e = tw::launch_check_finite(s->plan, flag, s->stream);
  int32_t h = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&h, flag, sizeof h, hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  (void)hipFree(flag);
  if (e != hipSuccess) return hip_fail(s, e, "check_finite");
  if (h) return fail(&s->err, TW_ENONFINITE, "non-finite value in the state");
  return TW_OK;
}

int tw_stream_copy(void* dst, const void* src, size_t bytes, void* hip_stream) {
  if (!dst || !src || bytes % 16 || ((uintptr_t)dst | (uintptr_t)src) % 16)
    return fail(nullptr, TW_EINVAL, "stream copy: 16-byte aligned pointers and a multiple of 16 bytes");
  const hipError_t e = tw::launch_stream_copy(dst, src, bytes, (hipStream_t)hip_stream);
  if (e != hipSuccess) return fail(nullptr, TW_EDEVICE, std::string("stream copy: ") + hipGetErrorString(e));
  return TW_OK;
}

int tw_check_finite_noalloc(tw_solver* s) {
  if (!s) return fail(nullptr, TW_EINVAL, "null handle");
  (void)hipSetDevice(s->cfg.device);
  int32_t h = 0;
  hipError_t e = tw::check_finite_scratch(s->plan, s->cfg.device, s->stream, &h);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) return hip_fail(s, e, "check_finite_noalloc");
  if (h) return fail(&s->err, TW_ENONFINITE, "non-finite value in the state");
  return TW_OK;
}

}  // extern "C"

// ------------------------------------------------------------ discrete adjoint
namespace {

// the fields of the adjoint state and checkpoints, and the coefficients with a gradient
std::vector<int> adj_fields(const tw_solver* s) {
  if (s->cfg.physics == TW_ELASTIC) return {TW_VX, TW_VY, TW_SXX, TW_SYY, TW_SXY};
  return {TW_VX, TW_VY, TW_P};
}
std::vector<int> adj_coefs(const tw_solver* s) {
  if (s->cfg.physics == TW_ELASTIC) return {TW_C_BINV_VX, TW_C_BINV_VY, TW_C_MU, TW_C_LAMBDA, TW_C_TWOMU};
  return {TW_C_KAPPA, TW_C_BINV_VX, TW_C_BINV_VY};
}
// the field whose grid (and pitch) a coefficient shares
int coef_field(int k) {
  switch (k) {
    case TW_C_BINV_VX: return TW_VX;
    case TW_C_BINV_VY: return TW_VY;
    case TW_C_MU: return TW_SXY;
    case TW_C_KAPPA: return TW_P;
    default: return TW_SXX;
  }
}
// line length of side `side`'s boundary points of velocity f (across the axis)
int side_len(const tw::Plan& p, int side, int f) { return side < 2 ? p.f[f].cols : p.f[f].rows; }

// forward taps along an axis of M cells (the per-point kernels' ddx/ddy):
// out_t 1 = N output (M + 1 points, partner M), 0 = M output (partner N)
void fwd_taps(const tw::Tables& t, int M, bool periodic, int out_t, std::vector<int>& nt,
              std::vector<std::array<int, 5>>& idx, std::vector<std::array<double, 5>>& cf) {
  const int n = out_t ? M + 1 : M;
  nt.assign(n, 0);
  idx.assign(n, {});
  cf.assign(n, {});
  const int sh0 = out_t ? -2 : -1, nb = out_t ? 4 : 3, np = out_t ? M : M + 1;
  for (int i = 0; i < n; i++) {
    if (periodic) {
      nt[i] = 4;
      for (int k = 0; k < 4; k++) {
        idx[i][k] = ((i + sh0 + k) % M + M) % M;
        cf[i][k] = t.st[k];
      }
    } else if (i < nb) {
      nt[i] = 5;
      for (int w = 0; w < 5; w++) {
        idx[i][w] = w;
        cf[i][w] = out_t ? t.DN_L[i][w] : t.DM_L[i][w];
      }
    } else if (i >= n - nb) {
      const int r = i - (n - nb);
      nt[i] = 5;
      for (int w = 0; w < 5; w++) {
        idx[i][w] = np - 5 + w;
        cf[i][w] = out_t ? t.DN_R[r][w] : t.DM_R[r][w];
      }
    } else {
      nt[i] = 4;
      for (int k = 0; k < 4; k++) {
        idx[i][k] = i + sh0 + k;
        cf[i][k] = t.st[k];
      }
    }
  }
}

// transpose of the taps, uploaded as a TapList (coefficients in the
// arithmetic type)
int upload_transpose(tw_solver* s, const tw::Tables& t, int M, bool periodic, int out_t, tw::TapList* L) {
  std::vector<int> nt;
  std::vector<std::array<int, 5>> idx;
  std::vector<std::array<double, 5>> cf;
  fwd_taps(t, M, periodic, out_t, nt, idx, cf);
  const int np = out_t ? M : M + 1;
  std::vector<int32_t> tn(np, 0), ti((size_t)np * tw::TMAX, 0);
  std::vector<double> tc((size_t)np * tw::TMAX, 0.0);
  for (int i = 0; i < (int)nt.size(); i++)
    for (int k = 0; k < nt[i]; k++) {
      const int j = idx[i][k];
      if (tn[j] >= tw::TMAX) return fail(&s->err, TW_EINVAL, "adjoint tap list overflow");
      ti[(size_t)j * tw::TMAX + tn[j]] = i;
      tc[(size_t)j * tw::TMAX + tn[j]] = cf[i][k];
      tn[j]++;
    }
  const size_t as = arith_size(s->cfg.precision);
  int e;
  if ((e = dalloc(s, (void**)&L->n, tn.size() * 4))) return e;
  if ((e = dalloc(s, (void**)&L->idx, ti.size() * 4))) return e;
  if ((e = dalloc(s, &L->cf, tc.size() * as))) return e;
  HIPCHK(s, hipMemcpy(L->n, tn.data(), tn.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(s, hipMemcpy(L->idx, ti.data(), ti.size() * 4, hipMemcpyHostToDevice));
  if (as == 8) {
    HIPCHK(s, hipMemcpy(L->cf, tc.data(), tc.size() * 8, hipMemcpyHostToDevice));
  } else {
    std::vector<float> f(tc.begin(), tc.end());
    HIPCHK(s, hipMemcpy(L->cf, f.data(), f.size() * 4, hipMemcpyHostToDevice));
  }
  return TW_OK;
}

int adj_check(tw_solver* s) {
  const tw_config* c = &s->cfg;
  if (c->precision != TW_FP64 && c->precision != TW_FP32) return fail(&s->err, TW_EINVAL, "adjoint: FP64 or FP32 storage only");
  if (c->cpst != TW_CPST_0) return fail(&s->err, TW_EINVAL, "adjoint: no compensation (cpst 0) only");
  return TW_OK;
}

int adj_setup(tw_solver* s) {
  if (s->adj_taps) return TW_OK;
  Plan& p = s->plan;
  const size_t as = arith_size(s->cfg.precision);
  int e;
  for (int f : adj_fields(s)) {
    const tw::DevField& d = p.f[f];
    if ((e = dalloc(s, &s->adj.mu[f], (size_t)d.rows * d.pitch * as))) return e;
  }
  for (int k : adj_coefs(s)) {
    const tw::DevField& d = p.f[coef_field(k)];
    if ((e = dalloc(s, (void**)&s->adj.g[k], (size_t)d.rows * d.pitch * 8))) return e;
  }
  const bool el = s->cfg.physics == TW_ELASTIC;
  for (int side = 0; side < 4; side++) {
    // normal velocity of each side (vx on x sides, vy on y sides); elastic:
    // the tangential one too (vy on x sides, vx on y sides)
    const int fn = side < 2 ? TW_VX : TW_VY, ft = side < 2 ? TW_VY : TW_VX;
    if ((e = dalloc(s, (void**)&s->adj.gz[side], (size_t)side_len(p, side, fn) * 8))) return e;
    if (!el) continue;
    if ((e = dalloc(s, (void**)&s->adj.gzt[side], (size_t)side_len(p, side, ft) * 8))) return e;
    if ((e = dalloc(s, &s->adj.dn[side], (size_t)side_len(p, side, fn) * as))) return e;
    if ((e = dalloc(s, &s->adj.dtg[side], (size_t)side_len(p, side, ft) * 3 * as))) return e;
  }
  if ((e = upload_transpose(s, p.t, p.mx, p.perx, 1, &s->adj.gt[0]))) return e;
  if ((e = upload_transpose(s, p.t, p.my, p.pery, 1, &s->adj.gt[1]))) return e;
  if ((e = upload_transpose(s, p.t, p.mx, p.perx, 0, &s->adj.dt[0]))) return e;
  if ((e = upload_transpose(s, p.t, p.my, p.pery, 0, &s->adj.dt[1]))) return e;
  s->adj_state_bytes = 0;
  for (int f : adj_fields(s)) s->adj_state_bytes += (size_t)p.f[f].rows * p.f[f].pitch * state_size(s->cfg.precision);
  s->adj_taps = true;
  return TW_OK;
}

// copy the live fields to / from a state slot
int state_copy(tw_solver* s, void* slot, bool save) {
  Plan& p = s->plan;
  char* q = (char*)slot;
  const size_t es = state_size(s->cfg.precision);
  for (int f : adj_fields(s)) {
    const size_t b = (size_t)p.f[f].rows * p.f[f].pitch * es;
    if (save)
      HIPCHK(s, hipMemcpyAsync(q, p.f[f].s, b, hipMemcpyDeviceToDevice, s->stream));
    else
      HIPCHK(s, hipMemcpyAsync(p.f[f].s, q, b, hipMemcpyDeviceToDevice, s->stream));
    q += b;
  }
  return TW_OK;
}

tw::StateView state_view(tw_solver* s, void* slot) {
  tw::StateView v;
  char* q = (char*)slot;
  const size_t es = state_size(s->cfg.precision);
  for (int f : adj_fields(s)) {
    v.s[f] = q;
    q += (size_t)s->plan.f[f].rows * s->plan.f[f].pitch * es;
  }
  return v;
}

int grow(tw_solver* s, void** buf, int* cap, int want, size_t unit) {
  if (*cap >= want) return TW_OK;
  if (*buf) {
    (void)hipStreamSynchronize(s->stream);
    dfree(s, buf);
  }
  const int e = dalloc(s, buf, unit * (size_t)want);
  if (e) return e;
  *cap = want;
  return TW_OK;
}

}  // namespace

int tw_adjoint_forward(tw_solver* s, int32_t nsteps, int32_t interval) {
  if (!s || nsteps <= 0) return fail(s ? &s->err : nullptr, TW_EINVAL, "bad argument");
  (void)hipSetDevice(s->cfg.device);
  int e;
  if ((e = adj_check(s)) || (e = adj_setup(s))) return e;
  if (interval <= 0) interval = (int)std::ceil(std::sqrt((double)nsteps));
  if (interval > nsteps) interval = nsteps;
  const int nck = (nsteps + interval - 1) / interval;
  const size_t sb = s->adj_state_bytes;
  if ((e = grow(s, &s->adj_ckpt, &s->adj_cap_ckpt, nck, sb))) return e;
  if ((e = grow(s, &s->adj_seg, &s->adj_cap_seg, interval + 1, sb))) return e;
  if (!s->adj_final && (e = dalloc(s, &s->adj_final, sb))) return e;
  s->adj_it0 = s->it;
  s->adj_n = nsteps;
  s->adj_I = interval;
  s->adj_nckpt = nck;
  for (int k = 0; k < nsteps; k++) {
    if (k % interval == 0 && (e = state_copy(s, (char*)s->adj_ckpt + sb * (size_t)(k / interval), true))) return e;
    if ((e = tw_step(s, 1))) return e;
  }
  if ((e = state_copy(s, s->adj_final, true))) return e;
  s->adj_done = true;
  return TW_OK;
}

namespace {

// z = sqrt(modulus / b) at a characteristic side's boundary point (the
// impedance upload_z sets): dz/dmodulus = z / (2 modulus), dz/db = -z / (2 b).
// Adds g_z dz/d(coefficient) to the coefficient gradients for the normal
// velocity (modulus kappa, or lambda + 2 mu of the adjacent MM cell) and, in
// the elastic case, the tangential one (mu of the adjacent NN node).
void impedance_chain(tw_solver* s, const std::vector<double>* gz, const std::vector<double>* gzt,
                     std::vector<double>* g) {
  const tw_config* c = &s->cfg;
  const bool el = c->physics == TW_ELASTIC;
  const int Mx = c->mx, My = c->my;
  for (int side = 0; side < 4; side++) {
    if (c->bc[side] != 'C') continue;
    const int axis = side / 2, right = side % 2;
    for (int f = TW_VX; f <= TW_VY; f++) {
      const bool normal = (axis == 0) == (f == TW_VX);
      const std::vector<double>& gs = normal ? gz[side] : gzt[side];
      if (gs.empty()) continue;
      const tw::DevField& F = s->plan.f[f];
      const int kb = f == TW_VX ? TW_C_BINV_VX : TW_C_BINV_VY;
      const std::vector<double>& b = s->hcoef[kb];
      for (int j = 0; j < (int)gs.size(); j++) {
        const int a = axis == 0 ? (right ? F.rows - 1 : 0) : (right ? F.cols - 1 : 0);
        const size_t qv = axis == 0 ? (size_t)a * F.cols + j : (size_t)j * F.cols + a;
        if (normal) {
          const int am = axis == 0 ? (right ? Mx - 1 : 0) : (right ? My - 1 : 0);
          const size_t qm = axis == 0 ? (size_t)am * My + j : (size_t)j * My + am;
          const double mod = el ? s->hcoef[TW_C_LAMBDA][qm] + s->hcoef[TW_C_TWOMU][qm] : s->hcoef[TW_C_KAPPA][qm];
          const double z = std::sqrt(mod / b[qv]), d = gs[j] * z / (2 * mod);
          if (el) {
            g[TW_C_LAMBDA][qm] += d;
            g[TW_C_TWOMU][qm] += d;
          } else {
            g[TW_C_KAPPA][qm] += d;
          }
          g[kb][qv] -= gs[j] * z / (2 * b[qv]);
        } else {
          const int an = axis == 0 ? (right ? Mx : 0) : (right ? My : 0);
          const size_t qn = axis == 0 ? (size_t)an * (My + 1) + j : (size_t)j * (My + 1) + an;
          const double mu = s->hcoef[TW_C_MU][qn];
          const double z = std::sqrt(mu / b[qv]);
          g[TW_C_MU][qn] += gs[j] * z / (2 * mu);
          g[kb][qv] -= gs[j] * z / (2 * b[qv]);
        }
      }
    }
  }
}

// the backward sweep of tw_adjoint_run / tw_adjoint_run_coefs: out[k] for
// coefficient id k (null: skipped)
int adjoint_sweep(tw_solver* s, const double* rcv_adjoint, double* const* out) {
  (void)hipSetDevice(s->cfg.device);
  Plan& p = s->plan;
  const int n = s->adj_n, I = s->adj_I, it0 = s->adj_it0;
  const size_t sb = s->adj_state_bytes;
  const bool el = s->cfg.physics == TW_ELASTIC;
  int e;
  // receiver sources
  const size_t ninj = (size_t)(s->nrcv > 0 ? s->nrcv : 1) * n;
  if (s->adj_inj_n < ninj) {
    if (s->adj_inj) dfree(s, (void**)&s->adj_inj);
    if ((e = dalloc(s, (void**)&s->adj_inj, ninj * 8))) return e;
    s->adj_inj_n = ninj;
  }
  if (s->nrcv > 0)
    HIPCHK(s, hipMemcpyAsync(s->adj_inj, rcv_adjoint, (size_t)s->nrcv * n * 8, hipMemcpyHostToDevice, s->stream));
  s->adj.inj = s->adj_inj;
  s->adj.nt = n;
  // receivers that share a grid point: chains in list order for the injection
  const std::vector<tw::Receiver>* lists[2] = {&s->rcv_v, &s->rcv_s};
  for (int g = 0; g < 2; g++) {
    const std::vector<tw::Receiver>& L = *lists[g];
    const size_t m = L.size();
    std::vector<int32_t> lead(m + 1, 0), next(m + 1, -1);
    std::map<std::array<int32_t, 3>, int> last;  // point -> last receiver seen there
    for (size_t r = 0; r < m; r++) {
      const std::array<int32_t, 3> key{L[r].field, L[r].ix, L[r].iy};
      auto it = last.find(key);
      if (it == last.end()) {
        lead[r] = 1;
      } else {
        next[it->second] = (int32_t)r;
      }
      last[key] = (int)r;
    }
    dfree(s, (void**)&s->adj.lead[g]);
    dfree(s, (void**)&s->adj.next[g]);
    if ((e = dalloc(s, (void**)&s->adj.lead[g], (m + 1) * 4))) return e;
    if ((e = dalloc(s, (void**)&s->adj.next[g], (m + 1) * 4))) return e;
    HIPCHK(s, hipMemcpyAsync(s->adj.lead[g], lead.data(), (m + 1) * 4, hipMemcpyHostToDevice, s->stream));
    HIPCHK(s, hipMemcpyAsync(s->adj.next[g], next.data(), (m + 1) * 4, hipMemcpyHostToDevice, s->stream));
  }
  const size_t as = arith_size(s->cfg.precision);
  for (int f : adj_fields(s))
    HIPCHK(s, hipMemsetAsync(s->adj.mu[f], 0, (size_t)p.f[f].rows * p.f[f].pitch * as, s->stream));
  for (int k : adj_coefs(s)) {
    const tw::DevField& d = p.f[coef_field(k)];
    HIPCHK(s, hipMemsetAsync(s->adj.g[k], 0, (size_t)d.rows * d.pitch * 8, s->stream));
  }
  for (int side = 0; side < 4; side++) {
    HIPCHK(s, hipMemsetAsync(s->adj.gz[side], 0, (size_t)side_len(p, side, side < 2 ? TW_VX : TW_VY) * 8, s->stream));
    if (el)
      HIPCHK(s, hipMemsetAsync(s->adj.gzt[side], 0, (size_t)side_len(p, side, side < 2 ? TW_VY : TW_VX) * 8, s->stream));
  }
  for (int k = s->adj_nckpt - 1; k >= 0; k--) {
    const int b = k * I, en = std::min(n, b + I);
    // recompute the segment's states: slot t - b + 1 holds the state after step t
    if ((e = state_copy(s, (char*)s->adj_ckpt + sb * (size_t)k, false))) return e;
    s->it = it0 + b;
    s->ep_valid = false;
    if ((e = state_copy(s, s->adj_seg, true))) return e;
    s->quiet = true;
    for (int t = b; t < en && !e; t++) {
      e = tw_step(s, 1);
      if (!e) e = state_copy(s, (char*)s->adj_seg + sb * (size_t)(t - b + 1), true);
    }
    s->quiet = false;
    if (e) return e;
    for (int t = en - 1; t >= b; t--) {
      const tw::StateView cur = state_view(s, (char*)s->adj_seg + sb * (size_t)(t - b + 1));
      const tw::StateView prev = state_view(s, (char*)s->adj_seg + sb * (size_t)(t - b));
      HIPCHK(s, tw::launch_adjoint_step(p, s->adj, cur, prev, it0 + t, t, s->stream));
    }
  }
  // back to the state after the forward run
  if ((e = state_copy(s, s->adj_final, false))) return e;
  s->it = it0 + n;
  s->ep_valid = false;
  std::vector<double> g[6];
  for (int k : adj_coefs(s)) {
    const tw::DevField& d = p.f[coef_field(k)];
    g[k].resize((size_t)d.rows * d.cols);
    HIPCHK(s, hipMemcpy2DAsync(g[k].data(), (size_t)d.cols * 8, s->adj.g[k], (size_t)d.pitch * 8, (size_t)d.cols * 8,
                               d.rows, hipMemcpyDeviceToHost, s->stream));
  }
  std::vector<double> gz[4], gzt[4];
  for (int side = 0; side < 4; side++) {
    if (s->cfg.bc[side] != 'C') continue;
    gz[side].resize(side_len(p, side, side < 2 ? TW_VX : TW_VY));
    HIPCHK(s, hipMemcpyAsync(gz[side].data(), s->adj.gz[side], gz[side].size() * 8, hipMemcpyDeviceToHost, s->stream));
    if (!el) continue;
    gzt[side].resize(side_len(p, side, side < 2 ? TW_VY : TW_VX));
    HIPCHK(s, hipMemcpyAsync(gzt[side].data(), s->adj.gzt[side], gzt[side].size() * 8, hipMemcpyDeviceToHost,
                             s->stream));
  }
  HIPCHK(s, hipStreamSynchronize(s->stream));
  impedance_chain(s, gz, gzt, g);
  for (int k : adj_coefs(s))
    if (out[k]) memcpy(out[k], g[k].data(), g[k].size() * 8);
  return TW_OK;
}

int adjoint_ready(tw_solver* s, const double* rcv_adjoint) {
  if (!s || !rcv_adjoint) return fail(s ? &s->err : nullptr, TW_EINVAL, "null argument");
  if (!s->adj_done) return fail(&s->err, TW_ESTATE, "tw_adjoint_run needs a tw_adjoint_forward run first");
  return TW_OK;
}

}  // namespace

int tw_adjoint_run(tw_solver* s, const double* rcv_adjoint, double* grad_kappa, double* grad_binv_vx,
                   double* grad_binv_vy) {
  int e = adjoint_ready(s, rcv_adjoint);
  if (e) return e;
  if (s->cfg.physics != TW_ACOUSTIC) return fail(&s->err, TW_EINVAL, "tw_adjoint_run: acoustic; use tw_adjoint_run_coefs");
  double* out[6] = {};
  out[TW_C_KAPPA] = grad_kappa;
  out[TW_C_BINV_VX] = grad_binv_vx;
  out[TW_C_BINV_VY] = grad_binv_vy;
  return adjoint_sweep(s, rcv_adjoint, out);
}

int tw_adjoint_run_coefs(tw_solver* s, const double* rcv_adjoint, double* const* grads) {
  int e = adjoint_ready(s, rcv_adjoint);
  if (e) return e;
  if (!grads) return fail(&s->err, TW_EINVAL, "null argument");
  for (int k = 0; k < 6; k++)
    if (grads[k] && !coef_in(s->cfg.physics, k))
      return fail(&s->err, TW_EINVAL, "gradient of a coefficient this physics does not use");
  return adjoint_sweep(s, rcv_adjoint, grads);
}

// ------------------------------------------------------------ batched shots
struct tw_batch {
  std::vector<tw_solver*> sv;
  int device = 0;
  int rows = 16;  // chunk height (rows per block) of the core launches
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;  // around the last graph launch
  // cached graph of the last step range
  hipGraphExec_t exec = nullptr;
  int g_nsteps = -1, g_rows = -1;
  std::vector<int> g_it0;
  std::vector<uint64_t> g_gen;  // each solver's generation when the graph was built
  std::vector<char> g_ep;       // each solver's ep_valid before the graph was built
  void* items = nullptr;
  size_t items_cap = 0;
  // per step and half: item offsets and launch sizes
  struct Half {
    size_t core_off, band_off;
    int ncore, core_blocks, nband, band_blocks;
  };
  std::vector<Half> halves;
  std::string err;
};

namespace {

int bfail(tw_batch* b, int code, const std::string& m) { return fail(b ? &b->err : nullptr, code, m); }

#define BHIP(b, call)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) return bfail((b), TW_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

tw::Dyn batch_dyn(tw_solver* s, int it, int half) {
  tw::Dyn d{};
  d.it = it;
  d.energy = 0;
  const bool vsrc = s->src_grid == TW_GRID_MN || s->src_grid == TW_GRID_NM;
  if (s->src_grid >= 0 && (half == 0) == vsrc && it < (int)s->src_inc.size()) {
    d.src_on = 1;
    d.src_grid = s->src_grid;
    d.src_ix = s->src_ix;
    d.src_iy = s->src_iy;
    d.src_inc = s->src_inc[it];
  } else {
    d.src_grid = -1;
  }
  d.part = half == 0 ? s->plan.part_v : s->plan.part_s;
  d.rec = s->plan.rec;
  d.nt_max = s->plan.nt_max;
  return d;
}

// item tables of nsteps steps from each solver's current step, in device
// memory, and the graph that launches them.  Energies and receiver samples are
// recorded as tw_step records them: each launch's first block of a problem
// reduces that problem's energy partials of the previous launch and samples
// the receivers of the fields it finished; the stress energy of the start
// state (when not known) and the last step's pending work are launches of
// their own per problem, inside the graph.
int batch_build(tw_batch* b, int nsteps) {
  const int n = (int)b->sv.size();
  std::vector<tw::BatchCore> core;
  std::vector<tw::BatchBand> band;
  b->halves.clear();
  // the batch's tile height for the geometry; each solver's own is restored
  // below, so its single-problem steps keep their partial-sum layout
  std::vector<int> txc0(n), kver0(n), nparts(n), has_core(n);
  for (int k = 0; k < n; k++) {
    tw_solver* s = b->sv[k];
    txc0[k] = s->plan.txc;
    s->plan.txc = b->rows;
    kver0[k] = s->plan.kver;
    s->plan.kver = 2;  // the batch runs the streaming kernels for every problem
    nparts[k] = tw::half_parts(s->plan);
    tw::BatchCore c{};
    tw::BatchBand r{};
    int nc, nb;
    tw::batch_geometry(s->plan, &c, &nc, &r, &nb);
    has_core[k] = nc > 0;
  }
  auto restore_txc = [&]() {
    for (int k = 0; k < n; k++) {
      b->sv[k]->plan.txc = txc0[k];
      b->sv[k]->plan.kver = kver0[k];
    }
  };
  for (int k = 0; k < n; k++) {
    tw_solver* s = b->sv[k];
    const tw_config* c = &s->cfg;
    const bool records = (c->record_energy || s->nrcv > 0) && s->it < c->nt_max;
    if (records && !has_core[k]) {
      restore_txc();
      return bfail(b, TW_EINVAL, "batch: a problem without core rows cannot record energy or receivers");
    }
    // energy partials of the batch's launch geometry
    const size_t need = (size_t)std::max(nparts[k], tw::energy_parts(s->plan));
    if (c->record_energy && need > s->part_cap) {
      dfree(s, (void**)&s->plan.part_v);
      dfree(s, (void**)&s->plan.part_s);
      int e;
      if ((e = dalloc(s, (void**)&s->plan.part_v, need * 8)) || (e = dalloc(s, (void**)&s->plan.part_s, need * 8))) {
        restore_txc();
        return bfail(b, e, "batch: energy partial buffers: " + s->err);
      }
      s->part_cap = need;
      s->gen++;  // graphs captured with the old buffers (other batches) are rebuilt
    }
  }
  std::vector<tw::Pending> carry(n);
  std::vector<char> have(n, 0), ep0(n);
  for (int k = 0; k < n; k++) ep0[k] = b->sv[k]->ep_valid;
  for (int t = 0; t < nsteps; t++)
    for (int half = 0; half < 2; half++) {
      tw_batch::Half H{core.size(), band.size(), 0, 0, 0, 0};
      for (int k = 0; k < n; k++) {
        tw_solver* s = b->sv[k];
        const tw_config* cf = &s->cfg;
        const tw::Plan& pl = s->plan;
        const int it = s->it + t;
        const bool rec = cf->record_energy && it < cf->nt_max;
        tw::BatchCore c{};
        tw::BatchBand r{};
        int nc, nb;
        tw::batch_geometry(pl, &c, &nc, &r, &nb);
        c.sp = r.sp = pl.dstat;
        tw::Dyn d = batch_dyn(s, it, half);
        d.energy = rec;
        if (half == 0) {
          if (t == 0 && rec && !s->ep_valid) {  // stress energy of the start state, launched first
            carry[k] = tw::Pending();
            carry[k].part = pl.part_s;
            carry[k].nparts = tw::energy_parts(pl);
            carry[k].out0 = pl.ep0;
            carry[k].it0 = it;
            carry[k].scale = cf->energy_form == 1 ? 1.0 : 0.5;
            have[k] = 1;
          }
          if (have[k]) d.pend = carry[k];
        } else {
          if (rec) {
            d.pend.part = pl.part_v;
            d.pend.nparts = nparts[k];
            d.pend.out0 = pl.ek;
            d.pend.it0 = it;
            d.pend.scale = cf->energy_form == 1 ? 1.0 : 0.5;
          }
          if (it < cf->nt_max && !s->rcv_v.empty()) {
            d.pend.rcv = pl.rcv_v;
            d.pend.nrcv = (int32_t)s->rcv_v.size();
            d.pend.rcv_it = it;
          }
          carry[k] = stress_pending(s, it, rec);
          carry[k].nparts = rec ? nparts[k] : 0;
          have[k] = 1;
        }
        c.d = r.d = d;
        if (nc > 0) {
          c.b0 = H.core_blocks;
          H.core_blocks += nc;
          core.push_back(c);
          H.ncore++;
        }
        if (nb > 0) {
          r.b0 = H.band_blocks;
          H.band_blocks += nb;
          // the problem's core blocks do its pending work; band tiles never do
          r.d.pend = tw::Pending();
          band.push_back(r);
          H.nband++;
        }
      }
      b->halves.push_back(H);
    }
  const size_t cb = core.size() * sizeof(tw::BatchCore), bb = band.size() * sizeof(tw::BatchBand);
  if (b->items_cap < cb + bb) {
    if (b->items) (void)hipFree(b->items);
    b->items = nullptr;
    if (hipMalloc(&b->items, cb + bb) != hipSuccess) {
      restore_txc();
      return bfail(b, TW_ENOMEM, "batch: item tables");
    }
    b->items_cap = cb + bb;
  }
  hipError_t he = hipMemcpyAsync(b->items, core.data(), cb, hipMemcpyHostToDevice, b->stream);
  if (he == hipSuccess)
    he = hipMemcpyAsync((char*)b->items + cb, band.data(), bb, hipMemcpyHostToDevice, b->stream);
  if (he == hipSuccess) he = hipStreamSynchronize(b->stream);
  if (he != hipSuccess) {
    restore_txc();
    return bfail(b, TW_EDEVICE, std::string("batch tables: ") + hipGetErrorString(he));
  }
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  b->exec = nullptr;
  const tw::BatchCore* dc = (const tw::BatchCore*)b->items;
  const tw::BatchBand* db = (const tw::BatchBand*)((char*)b->items + cb);
  const Plan& p0 = b->sv[0]->plan;
  hipGraph_t g = nullptr;
  he = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
  for (int k = 0; k < n && he == hipSuccess; k++) {
    tw_solver* s = b->sv[k];
    if (s->cfg.record_energy && s->it < s->cfg.nt_max && !s->ep_valid) he = tw::launch_stress_energy(s->plan, b->stream);
  }
  for (int i = 0; i < (int)b->halves.size() && he == hipSuccess; i++) {
    const tw_batch::Half& H = b->halves[i];
    he = tw::launch_batch_half(p0, i % 2, dc + H.core_off, H.ncore, H.core_blocks, db + H.band_off, H.nband,
                               H.band_blocks, b->stream);
  }
  for (int k = 0; k < n && he == hipSuccess; k++)
    if (have[k]) he = tw::launch_finalize(b->sv[k]->plan, carry[k], b->stream);
  if (he != hipSuccess) {
    (void)hipStreamEndCapture(b->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    restore_txc();
    return bfail(b, TW_EDEVICE, std::string("batched launch: ") + hipGetErrorString(he));
  }
  he = hipStreamEndCapture(b->stream, &g);
  if (he == hipSuccess) he = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  restore_txc();
  if (he != hipSuccess) return bfail(b, TW_EDEVICE, std::string("batch graph: ") + hipGetErrorString(he));
  b->g_nsteps = nsteps;
  b->g_rows = b->rows;
  b->g_it0.clear();
  b->g_gen.clear();
  b->g_ep.clear();
  for (int k = 0; k < n; k++) {
    b->g_it0.push_back(b->sv[k]->it);
    b->g_gen.push_back(b->sv[k]->gen);
    b->g_ep.push_back(ep0[k]);
  }
  return TW_OK;
}

bool batch_cached(const tw_batch* b, int nsteps) {
  if (!b->exec || b->g_nsteps != nsteps || b->g_rows != b->rows) return false;
  for (size_t k = 0; k < b->sv.size(); k++)
    if (b->g_it0[k] != b->sv[k]->it || b->g_gen[k] != b->sv[k]->gen || b->g_ep[k] != (char)b->sv[k]->ep_valid)
      return false;
  return true;
}

// the graph launch on the batch stream, ordered after each solver's queued
// work and before anything queued on the solvers afterwards
int batch_launch(tw_batch* b, int nsteps, hipEvent_t t0, hipEvent_t t1) {
  for (tw_solver* s : b->sv) {
    BHIP(b, hipEventRecord(b->done, s->stream));
    BHIP(b, hipStreamWaitEvent(b->stream, b->done, 0));
  }
  if (t0) BHIP(b, hipEventRecord(t0, b->stream));
  BHIP(b, hipGraphLaunch(b->exec, b->stream));
  if (t1) BHIP(b, hipEventRecord(t1, b->stream));
  BHIP(b, hipEventRecord(b->done, b->stream));
  for (tw_solver* s : b->sv) {
    BHIP(b, hipStreamWaitEvent(s->stream, b->done, 0));
    // the graph carried the stress energy of each recorded step into E_p0 of
    // the next, as tw_step does
    if (s->cfg.record_energy && s->it < s->cfg.nt_max) s->ep_valid = true;
    s->it += nsteps;
  }
  return TW_OK;
}

// device copies of every field array (state and companions) of the batch
int batch_state(tw_batch* b, std::vector<void*>& bufs, bool save) {
  size_t k = 0;
  for (tw_solver* s : b->sv) {
    for (int f = 0; f < tw::NF; f++) {
      const tw::DevField& d = s->plan.f[f];
      void* arr[3] = {d.s, d.r, d.c};
      for (int w = 0; w < 3; w++) {
        void* a = arr[w];
        if (!a) continue;
        const size_t bytes = (size_t)d.rows * d.pitch * array_size(&s->cfg, w);
        if (save) {
          if (k == bufs.size()) {
            void* q = nullptr;
            BHIP(b, hipMalloc(&q, bytes));
            bufs.push_back(q);
          }
          BHIP(b, hipMemcpyAsync(bufs[k], a, bytes, hipMemcpyDeviceToDevice, b->stream));
        } else {
          BHIP(b, hipMemcpyAsync(a, bufs[k], bytes, hipMemcpyDeviceToDevice, b->stream));
        }
        k++;
      }
    }
  }
  return TW_OK;
}

}  // namespace

int tw_batch_create(tw_solver* const* solvers, int32_t n, tw_batch** out) {
  if (!solvers || n <= 0 || !out) return fail(nullptr, TW_EINVAL, "bad argument");
  *out = nullptr;
  const tw_config& c0 = solvers[0]->cfg;
  for (int k = 0; k < n; k++) {
    const tw_solver* s = solvers[k];
    if (!s) return fail(nullptr, TW_EINVAL, "null solver");
    const tw_config& c = s->cfg;
    if (c.physics != c0.physics || c.precision != c0.precision || c.coef16 != c0.coef16 || c.cpst != c0.cpst ||
        c.device != c0.device || c.energy_form != c0.energy_form)
      return fail(nullptr, TW_EINVAL, "batch: solvers must share physics, precision, coef16, cpst, device, energy form");
    if (s->need_uhat) return fail(nullptr, TW_EINVAL, "batch: characteristic sides of elastic problems are not batched");
    if (s->plan.fused) return fail(nullptr, TW_EINVAL, "batch: fused-step solvers are not batched");
    if ((s->plan.kver == 1 && !s->auto_point) || c.cpst == TW_CPST_3C || c.cpst == TW_CPST_M3C || c.cpst == TW_CPST_6C)
      return fail(nullptr, TW_EINVAL, "batch: streaming kernels only (no 3C/-3C/6C)");
  }
  (void)hipSetDevice(c0.device);
  tw_batch* b = new tw_batch();
  b->sv.assign(solvers, solvers + n);
  b->device = c0.device;
  b->rows = solvers[0]->plan.txc;
  if (hipStreamCreate(&b->stream) != hipSuccess || hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
    tw_batch_destroy(b);
    return fail(nullptr, TW_EDEVICE, "batch: stream or event creation failed");
  }
  *out = b;
  return TW_OK;
}

int tw_batch_destroy(tw_batch* b) {
  if (!b) return TW_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->items) (void)hipFree(b->items);
  if (b->done) (void)hipEventDestroy(b->done);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return TW_OK;
}

const char* tw_batch_last_error(const tw_batch* b) { return b ? b->err.c_str() : g_err.c_str(); }

int tw_batch_set_rows(tw_batch* b, int32_t rows) {
  if (!b || rows <= 0) return bfail(b, TW_EINVAL, "bad argument");
  b->rows = rows;
  return TW_OK;
}

int tw_batch_step(tw_batch* b, int32_t nsteps) {
  if (!b || nsteps < 0) return bfail(b, TW_EINVAL, "bad argument");
  if (nsteps == 0) return TW_OK;
  (void)hipSetDevice(b->device);
  int e;
  if (!batch_cached(b, nsteps) && (e = batch_build(b, nsteps))) return e;
  return batch_launch(b, nsteps, b->ev0, b->ev1);
}

int tw_batch_elapsed(tw_batch* b, double* ms) {
  if (!b || !ms) return bfail(b, TW_EINVAL, "bad argument");
  (void)hipSetDevice(b->device);
  float t = 0.f;
  BHIP(b, hipEventSynchronize(b->ev1));
  BHIP(b, hipEventElapsedTime(&t, b->ev0, b->ev1));
  *ms = t;
  return TW_OK;
}

int tw_batch_autotune(tw_batch* b, int32_t trial_steps, int32_t* best_rows) {
  if (!b || trial_steps <= 0) return bfail(b, TW_EINVAL, "bad argument");
  (void)hipSetDevice(b->device);
  std::vector<void*> bufs;
  std::vector<int> it0;
  for (tw_solver* s : b->sv) {
    BHIP(b, hipStreamSynchronize(s->stream));
    it0.push_back(s->it);
  }
  int e = batch_state(b, bufs, true);
  hipEvent_t t0 = nullptr, t1 = nullptr;
  BHIP(b, hipEventCreate(&t0));
  BHIP(b, hipEventCreate(&t1));
  const int cand[] = {4, 8, 16, 32, 64};
  float best = 0.f;
  int best_r = -1;
  for (int r : cand) {
    if (e) break;
    b->rows = r;
    if ((e = batch_build(b, trial_steps))) break;
    if ((e = batch_launch(b, trial_steps, t0, t1))) break;
    float ms = 0.f;
    if (hipEventSynchronize(t1) != hipSuccess || hipEventElapsedTime(&ms, t0, t1) != hipSuccess) {
      e = bfail(b, TW_EDEVICE, "batch autotune timing");
      break;
    }
    if (best_r < 0 || ms < best) {
      best = ms;
      best_r = r;
    }
    if ((e = batch_state(b, bufs, false))) break;
    for (size_t k = 0; k < b->sv.size(); k++) b->sv[k]->it = it0[k];
  }
  (void)hipStreamSynchronize(b->stream);
  for (void* q : bufs) (void)hipFree(q);
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (e) return e;
  b->rows = best_r;
  if (best_rows) *best_rows = best_r;
  return TW_OK;
}
