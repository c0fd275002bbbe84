// tw_internal.h -- plan shared by the host library (tw_host.cpp) and the
// gfx950 kernels (tw_kernels.hip).  Plain pointers and sizes only.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/tw.h"

namespace tw {

enum { NF = 6 };

// Records msg as the thread's last error (tw_global_error) and returns code.
int host_fail(int code, const char* msg);

// One field in HBM: row-major [rows][pitch] elements of the storage type.
struct DevField {
  void* s = nullptr;  // state
  void* r = nullptr;  // R companion (3R/-3R/6R residual, or the 16-bit Kahan residual)
  void* c = nullptr;  // C companion (3C/-3C/6C)
  void* s2 = nullptr; // fused step: the other buffer of the state's ping-pong pair
  void* r2 = nullptr; // and of the R companion's
  int32_t rows = 0, cols = 0;
  int64_t pitch = 0;  // elements
};

struct DevCoef {
  void* p = nullptr;  // coefficient storage type (FP64, FP32, 16-bit, or 8-bit codes)
  void* rp = nullptr; // 8-bit codes: FP32 offset and step of each row's affine map
  int32_t rows = 0, cols = 0;
  int64_t pitch = 0;
};

// Operator tables, already rounded to the arithmetic type and held as double
// (exactly representable); the kernel side converts them once per launch.
struct Tables {
  double st[4];
  double DN_L[4][5], DN_R[4][5];  // N-type output (partner M): W_intr.DM_*
  double DM_L[3][5], DM_R[3][5];  // M-type output (partner N): W_intr.DN_*
  double ainvN_L, ainvN_R;
  double ainvM_L[3], ainvM_R[3];
  double projM_L[3], projM_R[3];
  double hN[4], hM[3];
};

struct Receiver {
  int32_t field, ix, iy, row;  // row: index into the record array
};

// Work the first block of a launch does for the previous launch: reduce its
// energy partials in a fixed order and gather the receivers of the fields it
// finished (those fields are not written by the current launch).
struct Pending {
  const double* part = nullptr;
  int32_t nparts = 0;
  double* out0 = nullptr;  // out0[it0] = sum * scale
  int32_t it0 = 0;
  double* out1 = nullptr;  // out1[it1] = same value (stress energy carried to E_p0)
  int32_t it1 = 0;
  double scale = 1.0;       // 1/2 for the H-norm form (grid_energy.cpp:215)
  const Receiver* rcv = nullptr;
  int32_t nrcv = 0;
  int32_t rcv_it = 0;
};

// Decode of an 8- or 16-bit material coefficient: code q of a row whose
// affine map has FP32 offset and step, off + q * step with one rounding (a
// fused multiply-add, correctly rounded on the host's libm and on the GPU
// alike, so both sides decode to the same FP32 value).
__host__ __device__ inline float c8_dec(unsigned q, float off, float step) {
  return __builtin_fmaf((float)q, step, off);
}

// Row pitch in elements: columns rounded up to 128, so every row of every
// array (16-bit state, and the one-byte residual codes) starts on a 128-byte
// cache line; pad adds elements (TW_PITCH_PAD, measurements only).
inline int64_t pitch_of(int cols, int pad) { return ((int64_t)cols + 127) / 128 * 128 + pad; }

// Rows per chunk of the streaming kernels: each chunk re-reads 3 window rows
// of the x partners, longer chunks leave fewer blocks to balance the launch.
// Measured on the 8192^2 elastic workload over chunks of 8-32 rows: 16-bit
// storage with a residual array is fastest at 13-14 rows, FP32 at 20-24,
// 16-bit without a residual at 16.
inline int32_t default_chunk_rows(int precision, int cpst) {
  const bool sixteen = precision == TW_BF16 || precision == TW_FP16;
  if (sixteen) return (cpst == TW_CPST_K || cpst == TW_CPST_K8 || cpst == TW_CPST_K8S) ? 14 : 16;
  return precision == TW_FP32 ? 20 : 16;
}

struct Plan {
  int32_t physics;
  int32_t precision;
  int32_t coef16;
  int32_t cpst;
  int32_t mx, my;
  int32_t perx, pery;
  char bc[4];
  int32_t energy_form;
  int32_t kver = 2;   // 1: per-point kernels, 2: streaming kernels, 3: fused step
  int32_t fused = 0;  // fused ping-pong step in use
  int32_t parity = 0; // which buffer of each ping-pong pair is current
  void* dstat = nullptr;  // device copy of the static kernel arguments (in place, current buffers)
  void* dstat_ip[2] = {}; // in-place arguments for each parity
  void* dstat_pp[2] = {}; // fused-step arguments (read current, write other) for each parity
  int32_t txc = 16;   // rows per chunk of the streaming kernels (default_chunk_rows; TW_TXC overrides)
  int32_t merge_bands = 1;  // band tiles as extra blocks of the core launches (TW_MERGE_BANDS=0: own launch)
  // timing of this step's core launches: start/stop of velocity, stress (null: off)
  hipEvent_t prof[4] = {};
  double dx;
  Tables t;
  DevField f[NF];
  DevCoef k[NF];
  // characteristic SAT: impedance per side, arithmetic type, for VX (0) and VY (1)
  void* z[2][4];
  // boundary extrapolation of M-type velocities (vy on x sides, vx on y sides)
  void* uhat[2][4];
  // records
  double* rec = nullptr;  // [nrcv][nt_max]
  int32_t nt_max = 0;
  double* ek = nullptr;
  double* ep0 = nullptr;
  double* ep1 = nullptr;
  double* part_v = nullptr;  // energy partials of the velocity launch
  double* part_s = nullptr;  // energy partials of the stress launch
  int32_t nparts = 0;
  Receiver* rcv_v = nullptr;  // receivers on velocity fields
  int32_t nrcv_v = 0;
  Receiver* rcv_s = nullptr;  // receivers on stress / pressure fields
  int32_t nrcv_s = 0;
};

struct StepCtl {
  int32_t it = 0;
  int32_t energy = 0;  // record energy this step
  // source
  int32_t src_grid = -1, src_ix = 0, src_iy = 0;
  double src_inc = 0.0;
  int32_t src_on = 0;
  Pending pend;
};

// Discrete adjoint of the time step (device buffers owned by the solver).
// Transposed tap lists along an axis: for a partner index j, the outputs i
// whose stencil reads j and the coefficient (arithmetic type, dt/dx included).
constexpr int TMAX = 8;
struct TapList {
  int32_t* n = nullptr;    // [len] entries per partner index
  int32_t* idx = nullptr;  // [len][TMAX]
  void* cf = nullptr;      // [len][TMAX] coefficients
};
struct AdjPlan {
  void* mu[NF] = {};        // adjoint of the fields (arithmetic type, field pitch)
  double* g[6] = {};        // d misfit / d coefficient, by coefficient id (TW_C_*), on its grid (field pitch)
  double* gz[4] = {};       // d misfit / d impedance of the normal velocity along each characteristic side
  double* gzt[4] = {};      // elastic: the same for the tangential velocity
  // elastic: adjoint increments of the characteristic self-terms, computed
  // from the incoming adjoint before any is applied: one per line of the
  // normal velocity's boundary points, three per line of the tangential one's
  void* dn[4] = {};
  void* dtg[4] = {};
  TapList gt[2];            // transpose of the N-output derivative of M data (x, y)
  TapList dt[2];            // transpose of the M-output derivative of N data (x, y)
  const double* inj = nullptr;  // receiver adjoint sources [nrcv][nt]
  int32_t nt = 0;
  // per receiver list (velocity, stress): 1 for the first receiver of its grid
  // point, and the next receiver at the same point (-1: none)
  int32_t* lead[2] = {};
  int32_t* next[2] = {};
};
// forward fields of a saved state (the physics' storage arrays, field pitch)
struct StateView {
  void* s[NF] = {};
};

// Per-launch control.
struct Dyn {
  int32_t it, energy;
  int32_t src_on, src_grid, src_ix, src_iy;
  double src_inc;
  double* part;
  double* part2;  // fused step: stress-energy partials (part: velocity)
  Pending pend;
  double* rec;
  int32_t nt_max;
};

// Rectangles of the union index space covered by a per-point launch.  n = 0:
// the whole (mx+1) x (my+1) space on a 2-D grid.  Otherwise blockIdx.x runs
// over the tiles of the n rectangles in order (the SBP boundary band around
// the core region of the streaming kernels).  A tile is 256 points: tw[k]
// columns by 256 / tw[k] rows, tw = 4 or 8 for the narrow column bands so no
// thread of a block idles, 64 otherwise.
struct Rects {
  int32_t n;
  int32_t r0[4], r1[4], c0[4], c1[4];
  int32_t tw[4];   // tile width (columns) of each rectangle
  int32_t cum[5];  // first tile of each rectangle
  int32_t poff;    // offset of this launch's energy partials
};
__host__ __device__ inline int tile_width(int cols) { return cols <= 4 ? 4 : (cols <= 8 ? 8 : 64); }

// Core region of a half step: rows [rlo, rhi) and columns [clo, chi) of the
// union index space, away from SBP closures (all rows / columns along a
// periodic axis).  The boundary band around it is done by the per-point
// kernels launched over the band rectangles.
struct Core {
  int32_t rlo, rhi, clo, chi, txc;
};

// Batched shots: one problem's share of a batched half-step launch.
struct BatchCore {
  const void* sp;  // the problem's device Stat
  Dyn d;
  Core c;
  int32_t b0;      // first block of this problem in the launch
  int32_t nbx;     // strips
  int32_t nb;      // blocks
};
struct BatchBand {
  const void* sp;
  Dyn d;
  Rects r;
  int32_t b0;
  int32_t nb;
};

// launches (tw_kernels.hip, one build per storage type; the entry points that
// depend on it are routed by tw_dispatch.cpp)
void grid_dims(const Plan& p, dim3* grid, dim3* block);
int energy_parts(const Plan& p);  // partial sums written by launch_stress_energy
int half_parts(const Plan& p);    // partial sums written per half step
// bytes of the static kernel arguments; fills `out` (host memory) when non-null.
// pp: updated values go to the other buffers (s2, r2) instead of in place
size_t stat_fill(const Plan& p, void* out, int pp = 0);
// fused elastic step: eligible configuration, partial sums per half step,
// and the three launches of a step (0 band velocity, 1 fused core, 2 band stress)
bool fused_ok(const Plan& p);
int fused_parts(const Plan& p);
hipError_t launch_fused(const Plan& p, int what, const StepCtl& c, hipStream_t s);
hipError_t launch_half(const Plan& p, int half, const StepCtl& c, hipStream_t s);
hipError_t launch_uhat(const Plan& p, hipStream_t s);
hipError_t launch_stress_energy(const Plan& p, hipStream_t s);
hipError_t launch_finalize(const Plan& p, const Pending& pend, hipStream_t s);
hipError_t launch_check_finite(const Plan& p, int32_t* flag, hipStream_t s);
hipError_t launch_stream_copy(void* dst, const void* src, size_t bytes, hipStream_t s);
// allocation-free finiteness check (tw_noalloc.hip): device address of the
// per-device scratch array (resolved once per device, the device current),
// and the check with its flag there, copied to *host_flag on stream s
hipError_t scratch_address(int device, int32_t** out);
hipError_t check_finite_scratch(const Plan& p, int device, hipStream_t s, int32_t* host_flag);
hipError_t launch_batch_half(const Plan& p, int half, const BatchCore* core, int ncore, int core_blocks,
                             const BatchBand* band, int nband, int band_blocks, hipStream_t s);
// core and band share of one problem (d, sp and b0 are left to the caller)
void batch_geometry(const Plan& p, BatchCore* core, int* core_blocks, BatchBand* band, int* band_blocks);
// one backward step of the acoustic adjoint for forward step `it`: inject the
// receiver sources of step it, then apply the transposed step using the saved
// states after (cur) and before (prev) step it, accumulating the gradients
hipError_t launch_adjoint_step(const Plan& p, const AdjPlan& a, const StateView& cur, const StateView& prev, int it,
                               int it_rel, hipStream_t s);

}  // namespace tw
