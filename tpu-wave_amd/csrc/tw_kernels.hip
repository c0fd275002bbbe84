// tw_kernels.hip -- gfx950 kernels of the staggered SBP-SAT time step.
//
// One launch per half step.  A thread owns one (ix, iy) of the union index
// space (mx+1) x (my+1) and, for each field of the half step that has that
// point, does the whole per-point update in registers:
//   x derivative (interior taps, SBP closure rows or periodic wrap) * coef
//   + SAT along x, + y derivative * coef, + SAT along y,
//   compensated update, point source, FP64 energy term.
// This fuses the reference's interior_x / boundary_x / secure_bdry_x /
// periodic_y / update / apply_source / weighted_square launches
// (cuda/src/cuda/grid_kernel_x.cut, grid_kernel_y.cut, grid_update.cut,
// grid.cu, grid_energy.cu) into one pass over HBM per half step.  The per-point
// floating-point operation order is the reference host twin's
// (cuda/src/cpp/grid_drvt.tpp, grid_update.tpp), so FP64/FP32 results are
// bitwise those of the CPU restatement.  Built with -ffp-contract=off.
//
// Blocks are 64 x 4 threads: a wavefront covers 64 consecutive columns of one
// row, so closure-row selection along x is uniform per wavefront and loads are
// row-contiguous.  Energy partial sums are per block (fixed-order tree); the
// first block of the next launch reduces them in block order, which keeps the
// result deterministic without atomics or an extra launch.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <type_traits>

#include "tw_internal.h"

// This file is compiled once per field storage type, TW_PREC = 0..3 (FP64,
// FP32, BF16, FP16), so the four sets of template instantiations build in
// parallel.  Each build defines the precision-dispatched entry points with the
// suffix _p<TW_PREC> (TW_PUB); tw_dispatch.cpp routes the names declared in
// tw_internal.h to them.  Geometry helpers without a storage type are defined
// by the TW_PREC == 0 build only.
#ifndef TW_PREC
#error "compile with -DTW_PREC=<0..3>"
#endif
#define TW_PNAME_(n, p) n##_p##p
#define TW_PNAME(n, p) TW_PNAME_(n, p)
#define TW_PUB(n) TW_PNAME(n, TW_PREC)

// return F<ST, KT, CT>(args) for this unit's storage type (TW_PREC) and the plan's
// coefficient storage (FP32, the 16-bit type, or 8-bit codes)
#if TW_PREC == 0
#define TW_DISPATCH(p, F, ...) return F<double, double, double>(__VA_ARGS__)
#elif TW_PREC == 1
#define TW_DISPATCH(p, F, ...) return F<float, float, float>(__VA_ARGS__)
#elif TW_PREC == 2
#define TW_DISPATCH(p, F, ...)                                     \
  do {                                                             \
    if ((p).coef16 == 2) return F<bf16, c8, float>(__VA_ARGS__);   \
    if ((p).coef16 == 3) return F<bf16, c16, float>(__VA_ARGS__);  \
    if ((p).coef16) return F<bf16, bf16, float>(__VA_ARGS__);      \
    return F<bf16, float, float>(__VA_ARGS__);                     \
  } while (0)
#else
#define TW_DISPATCH(p, F, ...)                                             \
  do {                                                                     \
    if ((p).coef16 == 2) return F<_Float16, c8, float>(__VA_ARGS__);       \
    if ((p).coef16 == 3) return F<_Float16, c16, float>(__VA_ARGS__);      \
    if ((p).coef16) return F<_Float16, _Float16, float>(__VA_ARGS__);      \
    return F<_Float16, float, float>(__VA_ARGS__);                         \
  } while (0)
#endif

namespace tw {

// ------------------------------------------------------------ storage types
struct bf16 {
  uint16_t u;
};
// 8-bit material coefficient: a code q in [0, 255] of its row's affine map
// value = off + q * step (FP32, one fused multiply-add); c8f is a decoded value
struct c8 {
  uint8_t q;
};
// 16-bit material coefficient: the same with codes in [0, 65535]
struct c16 {
  uint16_t q;
};
struct c8f {
  float f;
};
// coefficient storage as codes of a per-row affine map
template <class KT>
constexpr bool is_code = std::is_same<KT, c8>::value || std::is_same<KT, c16>::value;

template <class ST>
struct St;
template <>
struct St<double> {
  using C = double;
  static __device__ __forceinline__ double ld(double v) { return v; }
  static __device__ __forceinline__ double st(double v) { return v; }
};
template <>
struct St<float> {
  using C = float;
  static __device__ __forceinline__ float ld(float v) { return v; }
  static __device__ __forceinline__ float st(float v) { return v; }
};
template <>
struct St<bf16> {
  using C = float;
  static __device__ __forceinline__ float ld(bf16 v) { return __uint_as_float((uint32_t)v.u << 16); }
  // round to nearest even, NaN kept quiet (same rule as the host conversion)
  // round to nearest even (gfx950 v_cvt_pk_bf16_f32; FP32 denormals are kept,
  // so every finite value rounds as the host's bit-level RNE does)
  static __device__ __forceinline__ bf16 st(float f) {
    bf16 r;
    r.u = __builtin_bit_cast(uint16_t, (__bf16)f);
    return r;
  }
};
template <>
struct St<c8> {
  using C = float;
  static __device__ __forceinline__ float ld(c8f v) { return v.f; }
};
template <>
struct St<c16> {
  using C = float;
  static __device__ __forceinline__ float ld(c8f v) { return v.f; }
};
template <>
struct St<_Float16> {
  using C = float;
  static __device__ __forceinline__ float ld(_Float16 v) { return (float)v; }
  static __device__ __forceinline__ _Float16 st(float f) { return (_Float16)f; }
};

template <class KT, class CT, class V>
__device__ __forceinline__ CT kld(V v) {
  return (CT)St<KT>::ld(v);
}

// ------------------------------------------------------------ 8-bit residual
// TW_CPST_K8 keeps the rounding residual of a 16-bit store in one byte per
// value: an integer code q in [-127, 127], residual = q * 2^(max(e, 1) - bias)
// with e the exponent field of the stored value, i.e. q units of ulp / 256
// (bias 142 for BF16, whose ulp is 2^(e - 134); 33 for FP16, ulp 2^(e - 25)).
// Both scalings are exact; the code is the residual in those units rounded to
// nearest even and clamped.  In registers a code travels in the storage
// type's 16 bits as a sign-extended integer.
template <class ST>
__device__ __forceinline__ int k8_shift(ST hi) {
  const uint32_t b = __builtin_bit_cast(uint16_t, hi);
  if constexpr (std::is_same_v<ST, bf16>) {
    const int e = (b >> 7) & 0xff;
    return (e > 1 ? e : 1) - 142;
  } else {
    const int e = (b >> 10) & 0x1f;
    return (e > 1 ? e : 1) - 33;
  }
}
template <class ST>
__device__ __forceinline__ float k8_dec(ST q, ST hi) {
  return __builtin_ldexpf((float)__builtin_bit_cast(int16_t, q), k8_shift(hi));
}
template <class ST>
__device__ __forceinline__ ST k8_enc(float x, ST hi) {
  const float q = __builtin_rintf(__builtin_ldexpf(x, -k8_shift(hi)));
  const float c = __builtin_fminf(__builtin_fmaxf(q, -127.f), 127.f);
  return __builtin_bit_cast(ST, (int16_t)(int)c);
}
// the 8-bit residual variants (K8, and K8S whose stencils read the residual too)
__host__ __device__ __forceinline__ bool is_k8(int cpst) { return cpst == TW_CPST_K8 || cpst == TW_CPST_K8S; }
// residual mode of a compensation variant: 1 the 16-bit residual, 2 the byte code
__host__ __device__ __forceinline__ int kmode_of(int cpst) {
  return cpst == TW_CPST_K ? 1 : (is_k8(cpst) ? 2 : 0);
}
// companion element q of a per-point access; byte codes widen to the register form
template <class ST>
__device__ __forceinline__ ST rd_res(const ST* r, int64_t q, int km) {
  if constexpr (sizeof(ST) == 2)
    if (km == 2) return __builtin_bit_cast(ST, (int16_t)((const int8_t*)r)[q]);
  return r[q];
}
// a value as the stencils, SAT terms and boundary values read it: the stored
// value, or (rf, K8S) the stored value plus its decoded residual in FP32
template <class ST, class CT>
__device__ __forceinline__ CT pval(const ST* s, const ST* r, int64_t q, int rf) {
  CT v = (CT)St<ST>::ld(s[q]);
  if constexpr (sizeof(ST) == 2)
    if (rf) v = v + (CT)k8_dec(rd_res(r, q, 2), s[q]);
  return v;
}
template <class ST>
__device__ __forceinline__ void wr_res(ST* r, int64_t q, ST v, int km) {
  if constexpr (sizeof(ST) == 2)
    if (km == 2) {
      ((int8_t*)r)[q] = (int8_t)__builtin_bit_cast(int16_t, v);
      return;
    }
  r[q] = v;
}

// ------------------------------------------------------------ typed plan
template <class CT>
struct TabT {
  CT st[4];
  CT DN_L[4][5], DN_R[4][5];
  CT DM_L[3][5], DM_R[3][5];
  CT ainvN_L, ainvN_R, ainvM_L[3], ainvM_R[3];
  CT projM_L[3], projM_R[3];
};

template <class ST>
struct F {
  ST* s;   // state read by the launch
  ST* r;
  ST* c;
  ST* so;  // where updated values go: s itself, or the other buffer of a
  ST* ro;  // ping-pong pair (the fused step reads one buffer, writes the other)
  int32_t rows, cols;
  int64_t pitch;
  __device__ __forceinline__ int64_t at(int ix, int iy) const { return (int64_t)ix * pitch + iy; }
};

template <class KT>
struct K {
  const KT* p;
  int64_t pitch;
  const float* rp;  // 8- or 16-bit codes: offset and step of each row (rp[2 r], rp[2 r + 1])
  __device__ __forceinline__ auto get(int ix, int iy) const {
    if constexpr (is_code<KT>)
      return c8f{c8_dec(p[(int64_t)ix * pitch + iy].q, rp[2 * ix], rp[2 * ix + 1])};
    else
      return p[(int64_t)ix * pitch + iy];
  }
};

// Solver description that does not change between launches.  The streaming
// kernels read it from a device copy (so out-of-line helpers can take its
// address); the per-point kernels receive it by value inside Args.
template <class ST, class KT, class CT>
struct Stat {
  F<ST> f[NF];
  K<KT> k[NF];
  TabT<CT> t;
  double hN[4], hM[3];
  double dx;
  int32_t mx, my, perx, pery;
  int32_t bc[4];  // 0 periodic, 1 free surface, 2 characteristic
  int32_t cpst, eform, physics;
  const CT* z[2][4];
  CT* uhat[2][4];
};


template <class ST, class KT, class CT>
struct Args : Stat<ST, KT, CT>, Dyn {};

template <class ST, class KT, class CT>
static Args<ST, KT, CT> make_args(const Plan& p, const StepCtl* c, double* part) {
  Args<ST, KT, CT> a;
  for (int i = 0; i < NF; i++) {
    a.f[i].s = (ST*)p.f[i].s;
    a.f[i].r = (ST*)p.f[i].r;
    a.f[i].c = (ST*)p.f[i].c;
    a.f[i].so = a.f[i].s;
    a.f[i].ro = a.f[i].r;
    a.f[i].rows = p.f[i].rows;
    a.f[i].cols = p.f[i].cols;
    a.f[i].pitch = p.f[i].pitch;
    a.k[i].p = (const KT*)p.k[i].p;
    a.k[i].pitch = p.k[i].pitch;
    a.k[i].rp = (const float*)p.k[i].rp;
  }
  for (int i = 0; i < 4; i++) a.t.st[i] = (CT)p.t.st[i];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 5; j++) {
      a.t.DN_L[i][j] = (CT)p.t.DN_L[i][j];
      a.t.DN_R[i][j] = (CT)p.t.DN_R[i][j];
    }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 5; j++) {
      a.t.DM_L[i][j] = (CT)p.t.DM_L[i][j];
      a.t.DM_R[i][j] = (CT)p.t.DM_R[i][j];
    }
  a.t.ainvN_L = (CT)p.t.ainvN_L;
  a.t.ainvN_R = (CT)p.t.ainvN_R;
  for (int i = 0; i < 3; i++) {
    a.t.ainvM_L[i] = (CT)p.t.ainvM_L[i];
    a.t.ainvM_R[i] = (CT)p.t.ainvM_R[i];
    a.t.projM_L[i] = (CT)p.t.projM_L[i];
    a.t.projM_R[i] = (CT)p.t.projM_R[i];
    a.hM[i] = p.t.hM[i];
  }
  for (int i = 0; i < 4; i++) a.hN[i] = p.t.hN[i];
  a.dx = p.dx;
  a.mx = p.mx;
  a.my = p.my;
  a.perx = p.perx;
  a.pery = p.pery;
  for (int i = 0; i < 4; i++) a.bc[i] = p.bc[i] == 'P' ? 0 : (p.bc[i] == 'F' ? 1 : 2);
  a.cpst = p.cpst;
  a.eform = p.energy_form;
  a.physics = p.physics;
  for (int v = 0; v < 2; v++)
    for (int s = 0; s < 4; s++) {
      a.z[v][s] = (const CT*)p.z[v][s];
      a.uhat[v][s] = (CT*)p.uhat[v][s];
    }
  a.rec = p.rec;
  a.nt_max = p.nt_max;
  a.part = part;
  a.part2 = nullptr;
  if (c) {
    a.it = c->it;
    a.energy = c->energy;
    a.src_on = c->src_on;
    a.src_grid = c->src_grid;
    a.src_ix = c->src_ix;
    a.src_iy = c->src_iy;
    a.src_inc = c->src_inc;
    a.pend = c->pend;
  } else {
    a.it = 0;
    a.energy = 0;
    a.src_on = 0;
    a.src_grid = -1;
    a.src_ix = a.src_iy = 0;
    a.src_inc = 0.0;
  }
  return a;
}

// ------------------------------------------------------------ derivatives
// One multiply-accumulate step a + x * c of the derivative sums and of their
// products with the material coefficient.  16-bit storage (FM) computes it as
// one fused multiply-add, as nvcc contracts the reference's device loops
// (cuda/src/cuda/grid_kernel_x.cut:38-40 v_d += S * c, drvt_to_rths);
// FP32 and FP64 storage keep the separate product and sum of the reference's
// host code, whose printed outputs they reproduce digit for digit.
template <bool FM, class CT>
__device__ __forceinline__ CT mac(CT a, CT x, CT c) {
  if constexpr (FM)
    return __builtin_fmaf(x, c, a);
  else
    return a + x * c;
}

// Derivative along x onto output row ix of an output with n rows, type out_t
// (1 = N, 0 = M), from partner P.  v = 0; v = v + S*c per tap in tap order.
template <class ST, class CT>
__device__ __forceinline__ CT ddx(const F<ST>& P, int out_t, int ix, int iy, int n, int periodic, int M,
                                  const TabT<CT>& T, int rf = 0) {
  constexpr bool FM = sizeof(ST) == 2;
  CT v = (CT)0;
  if (periodic) {
    int r = ix + (out_t ? -2 : -1);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      int rr = r + k;
      rr = rr < 0 ? rr + M : (rr >= M ? rr - M : rr);
      v = mac<FM>(v, pval<ST, CT>(P.s, P.r, P.at(rr, iy), rf), T.st[k]);
    }
    return v;
  }
  const int nb = out_t ? 4 : 3;
  if (ix < nb) {
#pragma unroll
    for (int w = 0; w < 5; w++) v = mac<FM>(v, pval<ST, CT>(P.s, P.r, P.at(w, iy), rf), (out_t ? T.DN_L[ix][w] : T.DM_L[ix][w]));
  } else if (ix >= n - nb) {
    const int r = ix - (n - nb);
    const int b = P.rows - 5;
#pragma unroll
    for (int w = 0; w < 5; w++) v = mac<FM>(v, pval<ST, CT>(P.s, P.r, P.at(b + w, iy), rf), (out_t ? T.DN_R[r][w] : T.DM_R[r][w]));
  } else {
    int r = ix + (out_t ? -2 : -1);
#pragma unroll
    for (int k = 0; k < 4; k++) v = mac<FM>(v, pval<ST, CT>(P.s, P.r, P.at(r + k, iy), rf), T.st[k]);
  }
  return v;
}

template <class ST, class CT>
__device__ __forceinline__ CT ddy(const F<ST>& P, int out_t, int ix, int iy, int n, int periodic, int M,
                                  const TabT<CT>& T, int rf = 0) {
  constexpr bool FM = sizeof(ST) == 2;
  const int64_t row = (int64_t)ix * P.pitch;
  CT v = (CT)0;
  if (periodic) {
    int c = iy + (out_t ? -2 : -1);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      int cc = c + k;
      cc = cc < 0 ? cc + M : (cc >= M ? cc - M : cc);
      v = mac<FM>(v, pval<ST, CT>(P.s, P.r, row + cc, rf), T.st[k]);
    }
    return v;
  }
  const int nb = out_t ? 4 : 3;
  if (iy < nb) {
#pragma unroll
    for (int w = 0; w < 5; w++) v = mac<FM>(v, pval<ST, CT>(P.s, P.r, row + w, rf), (out_t ? T.DN_L[iy][w] : T.DM_L[iy][w]));
  } else if (iy >= n - nb) {
    const int r = iy - (n - nb);
    const int b = P.cols - 5;
#pragma unroll
    for (int w = 0; w < 5; w++) v = mac<FM>(v, pval<ST, CT>(P.s, P.r, row + b + w, rf), (out_t ? T.DN_R[r][w] : T.DM_R[r][w]));
  } else {
    int c = iy + (out_t ? -2 : -1);
#pragma unroll
    for (int k = 0; k < 4; k++) v = mac<FM>(v, pval<ST, CT>(P.s, P.r, row + c + k, rf), T.st[k]);
  }
  return v;
}

// SAT term on a velocity point (secure_bdry, cuda/src/cpp/grid_drvt.tpp:469-695).
// axis 0: boundary rows, j = column; axis 1: boundary columns, j = row.
// own: the point's old value (u_hat for an N-type output); uh: precomputed
// extrapolation for an M-type output.  Returns the increment to add to acc,
// already multiplied by the coefficient, or leaves acc unchanged.
template <class ST, class KT, class CT>
__device__ __forceinline__ CT sat_add(const Stat<ST, KT, CT>& A, CT acc, int vf, int pf, int axis, int out_t, int ix,
                                      int iy, CT own, CT coef) {
  const F<ST>& P = A.f[pf];
  const F<ST>& O = A.f[vf];
  const int a = axis == 0 ? ix : iy;          // index along the axis
  const int j = axis == 0 ? iy : ix;          // index across
  const int n = axis == 0 ? O.rows : O.cols;  // own length along the axis
  const int np = axis == 0 ? P.rows : P.cols;
  const int alen = out_t ? 1 : 3;
  const int plen = out_t ? 3 : 1;
#pragma unroll
  for (int side = 0; side < 2; side++) {
    const int bc = A.bc[2 * axis + side];
    if (bc == 0) continue;
    int w;
    if (side == 0) {
      if (a >= alen) continue;
      w = a;
    } else {
      if (a < n - alen) continue;
      w = a - (n - alen);
    }
    CT send = (CT)0;
    for (int q = 0; q < plen; q++) {
      const int pa = side == 0 ? q : np - plen + q;
      const CT pw = out_t ? (side == 0 ? A.t.projM_L[q] : A.t.projM_R[q]) : (CT)1;
      const CT sv = pval<ST, CT>(P.s, P.r, axis == 0 ? P.at(pa, j) : P.at(j, pa), A.cpst == TW_CPST_K8S);
      send = send + pw * sv;
    }
    CT recv = (CT)0;
    if (bc == 2) {
      const CT u = out_t ? own : A.uhat[vf == TW_VX ? 0 : 1][2 * axis + side][j];
      const CT zu = A.z[vf == TW_VX ? 0 : 1][2 * axis + side][j] * u;
      recv = side == 0 ? zu : -zu;
    }
    const CT g = send - recv;
    const CT eta = side == 0 ? (CT)1 : (CT)-1;
    const CT ai = out_t ? (side == 0 ? A.t.ainvN_L : A.t.ainvN_R) : (side == 0 ? A.t.ainvM_L[w] : A.t.ainvM_R[w]);
    const CT vd = (eta * ai) * g;
    acc = acc + vd * coef;
  }
  return acc;
}

// ------------------------------------------------------------ update
template <class ST, class CT>
__device__ __forceinline__ CT acc_init(const F<ST>& X, int64_t q, int cpst) {
  if (cpst == TW_CPST_3R || cpst == TW_CPST_M3R || cpst == TW_CPST_6R || cpst == TW_CPST_K)
    return (CT)St<ST>::ld(X.r[q]);
  if constexpr (sizeof(ST) == 2)
    if (is_k8(cpst)) return (CT)k8_dec(rd_res(X.r, q, 2), X.s[q]);
  return (CT)0;
}

// Compensated update (cuda/src/cuda/grid_update.cut:62-124).  Writes the state
// (and companion) and returns the value the point now represents.
template <class ST, class CT>
__device__ __forceinline__ void update(const F<ST>& X, int64_t q, CT acc, int cpst, ST* snew, ST* rnew) {
  using S = St<ST>;
  if constexpr (sizeof(ST) >= 4) {
    const CT s = S::ld(*snew);  // old state (caller loads it)
    const CT r = acc;
    switch (cpst) {
      case TW_CPST_0:
        *snew = s + r;
        break;
      case TW_CPST_3R: {
        CT a = s + r;
        CT z = a - s;
        *rnew = r - z;
        *snew = a;
      } break;
      case TW_CPST_M3R: {
        CT Y = r;
        CT T = s + Y;
        CT u = T - s;
        u = Y - u;
        *rnew = u;
        *snew = T;
      } break;
      case TW_CPST_6R: {
        CT a = s + r;
        CT pa = a - r;
        CT pb = a - pa;
        CT da = s - pa;
        CT db = r - pb;
        *snew = a;
        *rnew = da + db;
      } break;
      case TW_CPST_3C: {
        CT Y = r + (CT)X.c[q];
        CT a = s + Y;
        CT z = a - s;
        X.c[q] = Y - z;
        *snew = a;
      } break;
      case TW_CPST_M3C: {
        CT Y = r + (CT)X.c[q];
        CT T = s + Y;
        CT u = T - s;
        u = Y - u;
        X.c[q] = u;
        *snew = T;
      } break;
      default: {  // TW_CPST_6C
        CT Y = r + (CT)X.c[q];
        CT a = s + Y;
        CT pa = a - Y;
        CT pb = a - pa;
        CT da = s - pa;
        CT db = Y - pb;
        *snew = a;
        X.c[q] = da + db;
      } break;
    }
  } else {
    const CT t = S::ld(*snew) + acc;
    const ST hi = S::st(t);
    *snew = hi;
    if (cpst == TW_CPST_K) *rnew = S::st(t - S::ld(hi));
    if (is_k8(cpst)) *rnew = k8_enc(t - S::ld(hi), hi);
  }
}

// source: S = (T)((double)S + inc) after the update (forward_simulation.cpp:57-114)
template <class ST, class CT>
__device__ __forceinline__ void add_source(ST* snew, ST* rnew, double inc, int kmode) {
  using S = St<ST>;
  if constexpr (sizeof(ST) >= 4) {
    *snew = (ST)((double)*snew + inc);
  } else {
    if (kmode == 2) {
      const double v = (double)S::ld(*snew) + (double)k8_dec(*rnew, *snew) + inc;
      const float vf = (float)v;
      const ST hi = S::st(vf);
      *snew = hi;
      *rnew = k8_enc(vf - S::ld(hi), hi);
    } else if (kmode) {
      const double v = (double)S::ld(*snew) + (double)S::ld(*rnew) + inc;
      const float vf = (float)v;
      const ST hi = S::st(vf);
      *snew = hi;
      *rnew = S::st(vf - S::ld(hi));
    } else {
      *snew = S::st((float)((double)S::ld(*snew) + inc));
    }
  }
}

template <class ST, class CT>
__device__ __forceinline__ CT value_of(ST s, ST r, int kmode) {
  CT v = (CT)St<ST>::ld(s);
  if constexpr (sizeof(ST) < 4) {
    if (kmode == 2) v = v + (CT)k8_dec(r, s);
    else if (kmode) v = v + (CT)St<ST>::ld(r);
  }
  return v;
}

// H-norm factor along an axis (adjust_parameters_energy_periodic,
// cuda/src/cpp/grid_energy.cpp:58-177)
template <class ST, class KT, class CT>
__device__ __forceinline__ double hw(const Stat<ST, KT, CT>& A, int axis, int t, int i, int n) {
  const int periodic = axis == 0 ? A.perx : A.pery;
  if (periodic) return (t && i == n - 1) ? 0.0 : A.dx;
  const int nb = t ? 4 : 3;
  if (i < nb) return t ? A.hN[i] : A.hM[i];
  if (i >= n - nb) return t ? A.hN[n - 1 - i] : A.hM[n - 1 - i];
  return A.dx;
}

// energy term of one scalar field point (grid_energy.cpp:180-216 / cpu_acoustic.cpp:327-344)
template <class ST, class KT, class CT>
__device__ __forceinline__ double energy_pt(const Stat<ST, KT, CT>& A, CT v, CT prm, int tx, int ty, int ix, int iy,
                                            int rows, int cols) {
  if (A.eform == 1) {
    double w = A.dx * A.dx / 2.;
    if (ty && (iy == 0 || iy == cols - 1)) w /= 2.;
    if (tx && (ix == 0 || ix == rows - 1)) w /= 2.;
    return (double)((CT)((CT)(v * prm) * v)) * w;
  }
  const double comp = 1. / (double)prm;
  const double w = (comp * hw(A, 0, tx, ix, rows)) * hw(A, 1, ty, iy, cols);
  const double s = (double)v;
  return s * w * s;
}

// ------------------------------------------------------------ reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// block sum, fixed order; result valid in thread 0
__device__ __forceinline__ double block_sum(double v, double* sh) {
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  const int lane = tid & 63, wid = tid >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (tid == 0) {
    const int nw = (blockDim.x * blockDim.y + 63) >> 6;
    for (int w = 0; w < nw; w++) r += sh[w];
  }
  return r;
}

// work left by the previous launch, done by block 0
template <class ST, class KT, class CT>
__device__ void do_pending(const Stat<ST, KT, CT>& A, const Dyn& D, double* sh) {
  const Pending& P = D.pend;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  const int nth = blockDim.x * blockDim.y;
  if (P.nparts > 0) {
    double v = 0.0;
    for (int i = tid; i < P.nparts; i += nth) v += P.part[i];
    v = block_sum(v, sh);
    if (tid == 0) {
      v = v * P.scale;
      if (P.out0) P.out0[P.it0] = v;
      if (P.out1) P.out1[P.it1] = v;
    }
  }
  const int kmode = kmode_of(A.cpst);
  for (int r = tid; r < P.nrcv; r += nth) {
    const Receiver rc = P.rcv[r];
    const F<ST>& X = A.f[rc.field];
    const int64_t q = X.at(rc.ix, rc.iy);
    ST rr = X.s[q];
    if (kmode) rr = rd_res(X.r, q, kmode);
    D.rec[(int64_t)rc.row * D.nt_max + P.rcv_it] = (double)value_of<ST, CT>(X.s[q], rr, kmode);
  }
}

// ------------------------------------------------------------ half steps


__device__ __forceinline__ bool map_point(const Rects& R, int bx, int by, int* ix, int* iy, int* pidx) {
  if (R.n == 0) {
    *iy = bx * blockDim.x + threadIdx.x;
    *ix = by * blockDim.y + threadIdx.y;
    *pidx = by * gridDim.x + bx;
    return true;
  }
  const int b = bx;
  int k = 0;
  while (k + 1 < R.n && b >= R.cum[k + 1]) k++;
  const int t = b - R.cum[k];
  const int tw = R.tw[k], th = 256 / tw;
  const int tc = (R.c1[k] - R.c0[k] + tw - 1) / tw;
  const int lin = threadIdx.y * blockDim.x + threadIdx.x;
  *ix = R.r0[k] + (t / tc) * th + lin / tw;
  *iy = R.c0[k] + (t % tc) * tw + lin % tw;
  *pidx = R.poff + b;
  return *ix < R.r1[k] && *iy < R.c1[k];
}

// Velocity half step: vx (NM) and vy (MN).
// one point of the velocity half step (the body of the per-point kernels); returns its energy term
template <class ST, class KT, class CT>
__device__ __forceinline__ double vel_pt_e(const Stat<ST, KT, CT>& A, const Dyn& D, int ix, int iy, bool inr) {
  const int el = A.physics == TW_ELASTIC;
  const int kmode = kmode_of(A.cpst);
  const int rf = A.cpst == TW_CPST_K8S;  // stencils read value + residual
  constexpr bool FM = sizeof(ST) == 2;
  double e = 0.0;
  // ---- vx on NM: d/dx of sxx (elastic) or p, + d/dy of sxy (elastic)
  {
    const F<ST>& X = A.f[TW_VX];
    if (inr && ix < X.rows && iy < X.cols) {
      const int64_t q = X.at(ix, iy);
      const CT b = kld<KT, CT>(A.k[TW_C_BINV_VX].get(ix, iy));
      const CT own = pval<ST, CT>(X.s, X.r, q, rf);
      CT acc = acc_init<ST, CT>(X, q, A.cpst);
      const int pxf = el ? TW_SXX : TW_P;
      CT d = ddx<ST, CT>(A.f[pxf], 1, ix, iy, X.rows, A.perx, A.mx, A.t, rf);
      acc = mac<FM>(acc, d, b);
      if (!A.perx) acc = sat_add(A, acc, TW_VX, pxf, 0, 1, ix, iy, own, b);
      if (el) {
        d = ddy<ST, CT>(A.f[TW_SXY], 0, ix, iy, X.cols, A.pery, A.my, A.t, rf);
        acc = mac<FM>(acc, d, b);
        if (!A.pery) acc = sat_add(A, acc, TW_VX, TW_SXY, 1, 0, ix, iy, own, b);
      }
      ST sn = X.s[q], rn = sn;
      if (kmode || A.cpst == TW_CPST_3R || A.cpst == TW_CPST_M3R || A.cpst == TW_CPST_6R) rn = rd_res(X.r, q, kmode);
      update<ST, CT>(X, q, acc, A.cpst, &sn, &rn);
      if (D.src_on && D.src_grid == TW_GRID_NM && ix == D.src_ix && iy == D.src_iy)
        add_source<ST, CT>(&sn, &rn, D.src_inc, kmode);
      X.so[q] = sn;
      if (kmode || A.cpst == TW_CPST_3R || A.cpst == TW_CPST_M3R || A.cpst == TW_CPST_6R) wr_res(X.ro, q, rn, kmode);
      if (D.energy) e += energy_pt(A, value_of<ST, CT>(sn, rn, kmode), b, 1, 0, ix, iy, X.rows, X.cols);
    }
  }
  // ---- vy on MN: d/dx of sxy (elastic) + d/dy of syy (elastic) or p
  {
    const F<ST>& X = A.f[TW_VY];
    if (inr && ix < X.rows && iy < X.cols) {
      const int64_t q = X.at(ix, iy);
      const CT b = kld<KT, CT>(A.k[TW_C_BINV_VY].get(ix, iy));
      const CT own = pval<ST, CT>(X.s, X.r, q, rf);
      CT acc = acc_init<ST, CT>(X, q, A.cpst);
      if (el) {
        CT d = ddx<ST, CT>(A.f[TW_SXY], 0, ix, iy, X.rows, A.perx, A.mx, A.t, rf);
        acc = mac<FM>(acc, d, b);
        if (!A.perx) acc = sat_add(A, acc, TW_VY, TW_SXY, 0, 0, ix, iy, own, b);
      }
      const int pyf = el ? TW_SYY : TW_P;
      CT d = ddy<ST, CT>(A.f[pyf], 1, ix, iy, X.cols, A.pery, A.my, A.t, rf);
      acc = mac<FM>(acc, d, b);
      if (!A.pery) acc = sat_add(A, acc, TW_VY, pyf, 1, 1, ix, iy, own, b);
      ST sn = X.s[q], rn = sn;
      if (kmode || A.cpst == TW_CPST_3R || A.cpst == TW_CPST_M3R || A.cpst == TW_CPST_6R) rn = rd_res(X.r, q, kmode);
      update<ST, CT>(X, q, acc, A.cpst, &sn, &rn);
      if (D.src_on && D.src_grid == TW_GRID_MN && ix == D.src_ix && iy == D.src_iy)
        add_source<ST, CT>(&sn, &rn, D.src_inc, kmode);
      X.so[q] = sn;
      if (kmode || A.cpst == TW_CPST_3R || A.cpst == TW_CPST_M3R || A.cpst == TW_CPST_6R) wr_res(X.ro, q, rn, kmode);
      if (D.energy) e += energy_pt(A, value_of<ST, CT>(sn, rn, kmode), b, 0, 1, ix, iy, X.rows, X.cols);
    }
  }
  return e;
}

template <class ST, class KT, class CT>
__device__ __forceinline__ void vel_pt(const Stat<ST, KT, CT>& A, const Dyn& D, const Rects& R, int bx, int by,
                                       double* sh) {
  if (bx == 0 && by == 0) do_pending<ST, KT, CT>(A, D, sh);
  int ix, iy, pidx;
  const bool inr = map_point(R, bx, by, &ix, &iy, &pidx);
  const double e = vel_pt_e<ST, KT, CT>(A, D, ix, iy, inr);
  if (D.energy) {
    const double bsum = block_sum(e, sh);
    if (threadIdx.x == 0 && threadIdx.y == 0) D.part[pidx] = bsum;
  }
}

template <class ST, class KT, class CT>
__global__ void __launch_bounds__(256) k_velocity(Args<ST, KT, CT> A, Rects R) {
  __shared__ double sh[8];
  vel_pt<ST, KT, CT>(A, A, R, blockIdx.x, blockIdx.y, sh);
}

// Stress half step: sxy (NN) and sxx, syy (MM), or p (MM, acoustic).
// one point of the stress half step (the body of the per-point kernels); returns its energy term
template <class ST, class KT, class CT>
__device__ __forceinline__ double str_pt_e(const Stat<ST, KT, CT>& A, const Dyn& D, int ix, int iy, bool inr) {
  const int kmode = kmode_of(A.cpst);
  const int rf = A.cpst == TW_CPST_K8S;  // stencils read value + residual
  const int rmode = kmode || A.cpst == TW_CPST_3R || A.cpst == TW_CPST_M3R || A.cpst == TW_CPST_6R;
  constexpr bool FM = sizeof(ST) == 2;
  double e = 0.0;
  if (A.physics == TW_ELASTIC) {
    {
      const F<ST>& X = A.f[TW_SXY];
      if (inr && ix < X.rows && iy < X.cols) {
        const int64_t q = X.at(ix, iy);
        const CT mu = kld<KT, CT>(A.k[TW_C_MU].get(ix, iy));
        CT acc = acc_init<ST, CT>(X, q, A.cpst);
        CT d = ddx<ST, CT>(A.f[TW_VY], 1, ix, iy, X.rows, A.perx, A.mx, A.t, rf);
        acc = mac<FM>(acc, d, mu);
        d = ddy<ST, CT>(A.f[TW_VX], 1, ix, iy, X.cols, A.pery, A.my, A.t, rf);
        acc = mac<FM>(acc, d, mu);
        ST sn = X.s[q], rn = sn;
        if (rmode) rn = rd_res(X.r, q, kmode);
        update<ST, CT>(X, q, acc, A.cpst, &sn, &rn);
        if (D.src_on && D.src_grid == TW_GRID_NN && ix == D.src_ix && iy == D.src_iy)
          add_source<ST, CT>(&sn, &rn, D.src_inc, kmode);
        X.so[q] = sn;
        if (rmode) wr_res(X.ro, q, rn, kmode);
        if (D.energy) e += energy_pt(A, value_of<ST, CT>(sn, rn, kmode), mu, 1, 1, ix, iy, X.rows, X.cols);
      }
    }
    {
      const F<ST>& X0 = A.f[TW_SXX];
      const F<ST>& X1 = A.f[TW_SYY];
      if (inr && ix < X0.rows && iy < X0.cols) {
        const int64_t q = X0.at(ix, iy);
        const CT lam = kld<KT, CT>(A.k[TW_C_LAMBDA].get(ix, iy));
        const CT tmu = kld<KT, CT>(A.k[TW_C_TWOMU].get(ix, iy));
        CT a0 = acc_init<ST, CT>(X0, q, A.cpst);
        CT a1 = acc_init<ST, CT>(X1, q, A.cpst);
        CT v = ddx<ST, CT>(A.f[TW_VX], 0, ix, iy, X0.rows, A.perx, A.mx, A.t, rf);
        a0 = mac<FM>(a0, v, lam);
        a1 = mac<FM>(a1, v, lam);
        a0 = mac<FM>(a0, v, tmu);
        CT w = ddy<ST, CT>(A.f[TW_VY], 0, ix, iy, X0.cols, A.pery, A.my, A.t, rf);
        a0 = mac<FM>(a0, w, lam);
        a1 = mac<FM>(a1, w, lam);
        a1 = mac<FM>(a1, w, tmu);
        ST s0 = X0.s[q], r0 = s0, s1 = X1.s[q], r1 = s1;
        if (rmode) {
          r0 = rd_res(X0.r, q, kmode);
          r1 = rd_res(X1.r, q, kmode);
        }
        update<ST, CT>(X0, q, a0, A.cpst, &s0, &r0);
        update<ST, CT>(X1, q, a1, A.cpst, &s1, &r1);
        if (D.src_on && D.src_grid == TW_GRID_MM && ix == D.src_ix && iy == D.src_iy) {
          add_source<ST, CT>(&s0, &r0, D.src_inc, kmode);
          add_source<ST, CT>(&s1, &r1, D.src_inc, kmode);
        }
        X0.so[q] = s0;
        X1.so[q] = s1;
        if (rmode) {
          wr_res(X0.ro, q, r0, kmode);
          wr_res(X1.ro, q, r1, kmode);
        }
        if (D.energy) {
          const double L = (double)lam, M2 = (double)tmu;
          double p1 = (L + M2) / (M2 * (2. * L + M2));
          double p2 = -(L * 2.) / (M2 * (2. * L + M2));
          const double hx = hw(A, 0, 0, ix, X0.rows), hy = hw(A, 1, 0, iy, X0.cols);
          p1 = (p1 * hx) * hy;
          p2 = (p2 * hx) * hy;
          const double sxx = (double)value_of<ST, CT>(s0, r0, kmode);
          const double syy = (double)value_of<ST, CT>(s1, r1, kmode);
          e += sxx * p1 * sxx + syy * p1 * syy + sxx * p2 * syy;
        }
      }
    }
  } else {
    const F<ST>& X = A.f[TW_P];
    if (inr && ix < X.rows && iy < X.cols) {
      const int64_t q = X.at(ix, iy);
      const CT kap = kld<KT, CT>(A.k[TW_C_KAPPA].get(ix, iy));
      CT acc = acc_init<ST, CT>(X, q, A.cpst);
      CT d = ddx<ST, CT>(A.f[TW_VX], 0, ix, iy, X.rows, A.perx, A.mx, A.t, rf);
      acc = mac<FM>(acc, d, kap);
      d = ddy<ST, CT>(A.f[TW_VY], 0, ix, iy, X.cols, A.pery, A.my, A.t, rf);
      acc = mac<FM>(acc, d, kap);
      ST sn = X.s[q], rn = sn;
      if (rmode) rn = rd_res(X.r, q, kmode);
      update<ST, CT>(X, q, acc, A.cpst, &sn, &rn);
      if (D.src_on && D.src_grid == TW_GRID_MM && ix == D.src_ix && iy == D.src_iy)
        add_source<ST, CT>(&sn, &rn, D.src_inc, kmode);
      X.so[q] = sn;
      if (rmode) wr_res(X.ro, q, rn, kmode);
      if (D.energy) {
        const CT v = value_of<ST, CT>(sn, rn, kmode);
        if (A.eform == 1) e += energy_pt(A, v, kap, 0, 0, ix, iy, X.rows, X.cols);
        else e += energy_pt(A, v, kap, 0, 0, ix, iy, X.rows, X.cols);
      }
    }
  }
  return e;
}

template <class ST, class KT, class CT>
__device__ __forceinline__ void str_pt(const Stat<ST, KT, CT>& A, const Dyn& D, const Rects& R, int bx, int by,
                                       double* sh) {
  if (bx == 0 && by == 0) do_pending<ST, KT, CT>(A, D, sh);
  int ix, iy, pidx;
  const bool inr = map_point(R, bx, by, &ix, &iy, &pidx);
  const double e = str_pt_e<ST, KT, CT>(A, D, ix, iy, inr);
  if (D.energy) {
    const double bsum = block_sum(e, sh);
    if (threadIdx.x == 0 && threadIdx.y == 0) D.part[pidx] = bsum;
  }
}

template <class ST, class KT, class CT>
__global__ void __launch_bounds__(256) k_stress(Args<ST, KT, CT> A, Rects R) {
  __shared__ double sh[8];
  str_pt<ST, KT, CT>(A, A, R, blockIdx.x, blockIdx.y, sh);
}

// A band tile (256 points) run by one 64-lane block of a core launch: four
// points per lane, in the tile order of map_point; the wave's energy sum goes
// to the tile's partial.  No pending work here: the core tile (0, 0) of the
// launch does it.
__device__ __forceinline__ bool map_tile_point(const Rects& R, int b, int lin, int* ix, int* iy) {
  int k = 0;
  while (k + 1 < R.n && b >= R.cum[k + 1]) k++;
  const int t = b - R.cum[k];
  const int tw = R.tw[k], th = 256 / tw;
  const int tc = (R.c1[k] - R.c0[k] + tw - 1) / tw;
  *ix = R.r0[k] + (t / tc) * th + lin / tw;
  *iy = R.c0[k] + (t % tc) * tw + lin % tw;
  return *ix < R.r1[k] && *iy < R.c1[k];
}

template <class ST, class KT, class CT, int HALF>
__device__ __forceinline__ void band_tile(const Stat<ST, KT, CT>& A, const Dyn& D, const Rects& R, int b) {
  double e = 0.0;
#pragma unroll 1
  for (int sub = 0; sub < 4; sub++) {
    int ix, iy;
    const bool inr = map_tile_point(R, b, sub * 64 + (int)threadIdx.x, &ix, &iy);
    e += HALF == 0 ? vel_pt_e<ST, KT, CT>(A, D, ix, iy, inr) : str_pt_e<ST, KT, CT>(A, D, ix, iy, inr);
  }
  if (D.energy) {
    const double t = wave_sum(e);
    if (threadIdx.x == 0) D.part[R.poff + b] = t;
  }
}

// Energy of the stress fields only (start of a run / after the state was set).
template <class ST, class KT, class CT>
__global__ void __launch_bounds__(256) k_stress_energy(Args<ST, KT, CT> A) {
  __shared__ double sh[8];
  const int iy = blockIdx.x * blockDim.x + threadIdx.x;
  const int ix = blockIdx.y * blockDim.y + threadIdx.y;
  const int kmode = kmode_of(A.cpst);
  double e = 0.0;
  if (A.physics == TW_ELASTIC) {
    const F<ST>& X = A.f[TW_SXY];
    if (ix < X.rows && iy < X.cols) {
      const int64_t q = X.at(ix, iy);
      const CT mu = kld<KT, CT>(A.k[TW_C_MU].get(ix, iy));
      ST r = X.s[q];
      if (kmode) r = rd_res(X.r, q, kmode);
      e += energy_pt(A, value_of<ST, CT>(X.s[q], r, kmode), mu, 1, 1, ix, iy, X.rows, X.cols);
    }
    const F<ST>& X0 = A.f[TW_SXX];
    const F<ST>& X1 = A.f[TW_SYY];
    if (ix < X0.rows && iy < X0.cols) {
      const int64_t q = X0.at(ix, iy);
      const double L = (double)kld<KT, CT>(A.k[TW_C_LAMBDA].get(ix, iy));
      const double M2 = (double)kld<KT, CT>(A.k[TW_C_TWOMU].get(ix, iy));
      double p1 = (L + M2) / (M2 * (2. * L + M2));
      double p2 = -(L * 2.) / (M2 * (2. * L + M2));
      const double hx = hw(A, 0, 0, ix, X0.rows), hy = hw(A, 1, 0, iy, X0.cols);
      p1 = (p1 * hx) * hy;
      p2 = (p2 * hx) * hy;
      ST r0 = X0.s[q], r1 = X1.s[q];
      if (kmode) {
        r0 = rd_res(X0.r, q, kmode);
        r1 = rd_res(X1.r, q, kmode);
      }
      const double sxx = (double)value_of<ST, CT>(X0.s[q], r0, kmode);
      const double syy = (double)value_of<ST, CT>(X1.s[q], r1, kmode);
      e += sxx * p1 * sxx + syy * p1 * syy + sxx * p2 * syy;
    }
  } else {
    const F<ST>& X = A.f[TW_P];
    if (ix < X.rows && iy < X.cols) {
      const int64_t q = X.at(ix, iy);
      const CT kap = kld<KT, CT>(A.k[TW_C_KAPPA].get(ix, iy));
      ST r = X.s[q];
      if (kmode) r = rd_res(X.r, q, kmode);
      e += energy_pt(A, value_of<ST, CT>(X.s[q], r, kmode), kap, 0, 0, ix, iy, X.rows, X.cols);
    }
  }
  const double bsum = block_sum(e, sh);
  if (threadIdx.x == 0 && threadIdx.y == 0) A.part[blockIdx.y * gridDim.x + blockIdx.x] = bsum;
}

// Boundary extrapolation of the M-type velocities for the characteristic SAT:
// vy (MN) on the x sides, vx (NM) on the y sides (projection p_M, namespace_forward.hpp:153-154).
template <class ST, class KT, class CT>
__global__ void __launch_bounds__(256) k_uhat(Args<ST, KT, CT> A) {
  const int rf = A.cpst == TW_CPST_K8S;
  const int j = blockIdx.x * blockDim.x * blockDim.y + threadIdx.y * blockDim.x + threadIdx.x;
  for (int side = 0; side < 4; side++) {
    const int axis = side >> 1, right = side & 1;
    const int vf = axis == 0 ? TW_VY : TW_VX;
    CT* out = A.uhat[vf == TW_VX ? 0 : 1][side];
    if (!out) continue;
    const F<ST>& X = A.f[vf];
    const int n_across = axis == 0 ? X.cols : X.rows;
    const int n_along = axis == 0 ? X.rows : X.cols;
    if (j >= n_across) continue;
    const CT* pm = right ? A.t.projM_R : A.t.projM_L;
    CT u = (CT)0;
    for (int w = 0; w < 3; w++) {
      const int a = right ? n_along - 3 + w : w;
      const CT v = pval<ST, CT>(X.s, X.r, axis == 0 ? X.at(a, j) : X.at(j, a), rf);
      u = u + pm[w] * v;
    }
    out[j] = u;
  }
}

template <class ST, class KT, class CT>
__global__ void __launch_bounds__(256) k_finalize(Args<ST, KT, CT> A) {
  __shared__ double sh[8];
  do_pending<ST, KT, CT>(A, A, sh);
}

template <class ST>
__global__ void __launch_bounds__(256) k_finite(F<ST> X, int32_t* flag) {
  const int iy = blockIdx.x * blockDim.x + threadIdx.x;
  const int ix = blockIdx.y * blockDim.y + threadIdx.y;
  if (ix < X.rows && iy < X.cols) {
    const double v = (double)St<ST>::ld(X.s[X.at(ix, iy)]);
    if (!(v - v == 0.0)) *flag = 1;
  }
}

// ============================================================ streaming kernels
// One wavefront (a 64-thread block) owns a strip of 256 columns (4 per lane)
// and walks down a chunk of rows.  Partner rows needed by the x stencil stay
// in a rolling register window, so each is read from HBM once per chunk; the
// y stencil reads a per-row LDS segment of the y partner (strip + 8-column
// halo each side, periodic wrap resolved when the segment is filled).  SBP
// closure rows read their five partner rows directly; closure columns use the
// five-tap boundary rows on the LDS segment.  The per-point arithmetic is the
// same sequence as the per-point kernels above.
constexpr int VW = 4;
constexpr int STRIP = 64 * VW;
constexpr int HALO = 8;
constexpr int SEG = STRIP + 2 * HALO;

template <class ST>
struct alignas(4 * sizeof(ST)) V4 {
  ST v[4];
};

// raw integer vectors of the same size, for loads/stores through typed
// address-space pointers (struct copies cannot go through them)
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
template <int B>
struct RawOf;
template <>
struct RawOf<1> {
  using e = uint8_t;
  using v = uint32_t;
};
template <>
struct RawOf<2> {
  using e = uint16_t;
  using v = u32x2;
};
template <>
struct RawOf<4> {
  using e = uint32_t;
  using v = u32x4;
};
template <>
struct RawOf<8> {
  using e = uint64_t;
  struct alignas(32) v {
    u32x4 a, b;
  };
};

template <class ST, class PT>
__device__ __forceinline__ V4<ST> ld_v4(PT p) {
  V4<ST> o;
  if constexpr (sizeof(ST) == 8) {
    typedef __attribute__((address_space(1))) const u32x4* Q;
    const u32x4 a = ((Q)p)[0], b = ((Q)p)[1];
    __builtin_memcpy(&o.v[0], &a, 16);
    __builtin_memcpy(&o.v[2], &b, 16);
  } else {
    typedef typename RawOf<sizeof(ST)>::v R;
    typedef __attribute__((address_space(1))) const R* Q;
    const R r = *(Q)p;
    __builtin_memcpy(&o, &r, sizeof o);
  }
  return o;
}

// Stores of new state and residual rows: written once per half step and not
// read again before the next one, so (below FP64) they are non-temporal
// streaming stores that do not keep the lines in the caches.
template <class ST, class PT>
__device__ __forceinline__ void st_v4(PT p, const V4<ST>& v) {
  if constexpr (sizeof(ST) == 8) {
    typedef __attribute__((address_space(1))) u32x4* Q;
    u32x4 a, b;
    __builtin_memcpy(&a, &v.v[0], 16);
    __builtin_memcpy(&b, &v.v[2], 16);
    ((Q)p)[0] = a;
    ((Q)p)[1] = b;
  } else {
    typedef typename RawOf<sizeof(ST)>::v R;
    typedef __attribute__((address_space(1))) R* Q;
    R r;
    __builtin_memcpy(&r, &v, sizeof r);
    __builtin_nontemporal_store(r, (Q)p);
  }
}

// Loads of rows that a half step reads exactly once (coefficients, and the
// old state and residual of the fields it updates): non-temporal below FP64,
// so they do not displace the partner rows that neighbouring chunks and
// strips read again (window warm-up rows, halo lines).
template <class ST, class PT>
__device__ __forceinline__ V4<ST> ld_v4s(PT p) {
  if constexpr (sizeof(ST) == 8) {
    return ld_v4<ST>(p);
  } else {
    V4<ST> o;
    typedef typename RawOf<sizeof(ST)>::v R;
    typedef __attribute__((address_space(1))) const R* Q;
    const R r = __builtin_nontemporal_load((Q)p);
    __builtin_memcpy(&o, &r, sizeof o);
    return o;
  }
}

// one element through a global pointer
template <class ST, class PT>
__device__ __forceinline__ ST ld_1(PT p) {
  typedef typename RawOf<sizeof(ST)>::e E;
  typedef __attribute__((address_space(1))) const E* Q;
  const E e = *(Q)p;
  ST v;
  __builtin_memcpy(&v, &e, sizeof v);
  return v;
}

template <class ST, class PT>
__device__ __forceinline__ void st_1(PT p, const ST& v) {
  typedef typename RawOf<sizeof(ST)>::e E;
  typedef __attribute__((address_space(1))) E* Q;
  E e;
  __builtin_memcpy(&e, &v, sizeof e);
  *(Q)p = e;
}

// columns c..c+3 of a row (c multiple of 4, row valid) -> arithmetic type
template <class ST, class CT, class PT>
__device__ __forceinline__ void ld4(PT rowp, int c, CT o[4]) {
  const V4<ST> t = ld_v4<ST>(rowp + c);
#pragma unroll
  for (int j = 0; j < 4; j++) o[j] = (CT)St<ST>::ld(t.v[j]);
}


// value of partner column v of a row, with periodic mapping / bounds
template <class ST, class CT>
__device__ __forceinline__ CT ldcol(const ST* rowp, int v, int periodic, int M, int ncols) {
  if (periodic) v = ((v % M) + M) % M;
  if (v < 0 || v >= ncols) return (CT)0;
  return (CT)St<ST>::ld(rowp[v]);
}

// fill the LDS segment of partner row `row` for virtual columns base-HALO .. base+STRIP+HALO
template <class ST, class CT>
__device__ __forceinline__ void seg_fill(CT* seg, const F<ST>& P, int row, int base, int lane, int periodic, int M,
                                         int rf = 0) {
  if (rf) {  // K8S: value + residual, element by element (edge strips only)
    const int64_t r0 = (int64_t)row * P.pitch;
    auto col = [&](int v) -> CT {
      if (periodic) v = ((v % M) + M) % M;
      if (v < 0 || v >= P.cols) return (CT)0;
      return pval<ST, CT>(P.s, P.r, r0 + v, 1);
    };
#pragma unroll
    for (int j = 0; j < 4; j++) seg[HALO + lane * VW + j] = col(base + lane * VW + j);
    if (lane < 2) {
#pragma unroll
      for (int j = 0; j < 4; j++) seg[lane * 4 + j] = col(base - HALO + lane * 4 + j);
    } else if (lane >= 62) {
#pragma unroll
      for (int j = 0; j < 4; j++) seg[HALO + STRIP + (lane - 62) * 4 + j] = col(base + STRIP + (lane - 62) * 4 + j);
    }
    return;
  }
  const ST* rowp = P.s + (int64_t)row * P.pitch;
  const int v0 = base + lane * VW;
  CT o[4];
  if (v0 >= 0 && v0 + 3 < (periodic ? M : P.cols)) {
    ld4<ST, CT>(rowp, v0, o);
  } else {
#pragma unroll
    for (int j = 0; j < 4; j++) o[j] = ldcol<ST, CT>(rowp, v0 + j, periodic, M, P.cols);
  }
#pragma unroll
  for (int j = 0; j < 4; j++) seg[HALO + lane * VW + j] = o[j];
  if (lane < 2) {
#pragma unroll
    for (int j = 0; j < 4; j++) seg[lane * 4 + j] = ldcol<ST, CT>(rowp, base - HALO + lane * 4 + j, periodic, M, P.cols);
  } else if (lane >= 62) {
#pragma unroll
    for (int j = 0; j < 4; j++)
      seg[HALO + STRIP + (lane - 62) * 4 + j] = ldcol<ST, CT>(rowp, base + STRIP + (lane - 62) * 4 + j, periodic, M, P.cols);
  }
}

// interior x derivative from the register window (rows in tap order)
template <bool FM, class CT>
__device__ __forceinline__ CT dx_win(const CT w[4][4], int j, const CT st[4]) {
  CT v = (CT)0;
#pragma unroll
  for (int k = 0; k < 4; k++) v = mac<FM>(v, w[k][j], st[k]);
  return v;
}

// interior / periodic y derivative from the LDS segment
template <bool FM, class CT>
__device__ __forceinline__ CT dy_win(const CT* seg, int base, int col, int out_t, const CT st[4]) {
  CT v = (CT)0;
  const int p = col - base + HALO + (out_t ? -2 : -1);
#pragma unroll
  for (int k = 0; k < 4; k++) v = mac<FM>(v, seg[p + k], st[k]);
  return v;
}

template <class T>
__device__ __forceinline__ void shift_win(T w[4][4]) {
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int j = 0; j < 4; j++) w[k][j] = w[k + 1][j];
}


// field parameters held in registers for the whole kernel
// global (address space 1) pointer: the field pointers come from a struct in
// memory, so without the cast the compiler would emit flat accesses
template <class T>
using gptr = __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gptr<T> as_global(T* p) {
  return (gptr<T>)p;
}

template <class ST>
struct FR {
  gptr<ST> s;
  gptr<ST> r;
  gptr<ST> so;
  gptr<ST> ro;
  int64_t pitch;
  int32_t rows, cols;
  __device__ __forceinline__ void load(const F<ST>& f) {
    s = as_global(f.s);
    r = as_global(f.r);
    so = as_global(f.so);
    ro = as_global(f.ro);
    pitch = f.pitch;
    rows = f.rows;
    cols = f.cols;
  }
  __device__ __forceinline__ gptr<ST> row(int i) const { return s + (int64_t)i * pitch; }
  __device__ __forceinline__ gptr<ST> rrow(int i) const { return r + (int64_t)i * pitch; }
  __device__ __forceinline__ gptr<ST> orow(int i) const { return so + (int64_t)i * pitch; }
  __device__ __forceinline__ gptr<ST> orrow(int i) const { return ro + (int64_t)i * pitch; }
  // residual rows of the K8 variant: one byte per value, same pitch in elements
  __device__ __forceinline__ gptr<int8_t> brow(int i) const { return (gptr<int8_t>)r + (int64_t)i * pitch; }
};

// K8S reads partner values with their residual: the byte codes of a lane's
// four values of a row (c a multiple of 4, the same addressing as the values)
template <class ST>
__device__ __forceinline__ uint32_t ld_c4(const FR<ST>& P, int row, int c) {
  typedef __attribute__((address_space(1))) const uint32_t* Q;
  return *(Q)(P.brow(row) + c);
}
// four stored values -> arithmetic type, plus (RF) their decoded residuals
template <class ST, class CT, bool RF>
__device__ __forceinline__ void cvt_rf(const V4<ST>& v, uint32_t w, CT* o) {
#pragma unroll
  for (int j = 0; j < 4; j++) {
    o[j] = (CT)St<ST>::ld(v.v[j]);
    if constexpr (RF && sizeof(ST) == 2)
      o[j] = o[j] + (CT)k8_dec(__builtin_bit_cast(ST, (int16_t)(int8_t)(w >> (8 * j))), v.v[j]);
  }
}
// columns c..c+3 of row `row` (valid) as the stencils read them
template <class ST, class CT, bool RF>
__device__ __forceinline__ void ld4f(const FR<ST>& P, int row, int c, CT* o) {
  const V4<ST> t = ld_v4<ST>(P.row(row) + c);
  uint32_t w = 0;
  if constexpr (RF) w = ld_c4<ST>(P, row, c);
  cvt_rf<ST, CT, RF>(t, w, o);
}

template <class ST, class CT, bool RF = false>
__device__ __forceinline__ void ldx_r(const FR<ST>& P, int row, int c, int periodic, int M, CT o[4]) {
  if (periodic) row = row < 0 ? row + M : (row >= M ? row - M : row);
  if (row < 0 || row >= P.rows || c >= P.cols) {
#pragma unroll
    for (int j = 0; j < 4; j++) o[j] = (CT)0;
    return;
  }
  ld4f<ST, CT, RF>(P, row, c, o);
}

template <class ST, class CT, bool RF = false>
__device__ __forceinline__ void seg_fill_r(CT* seg, const FR<ST>& P, int row, int base, int lane, int periodic, int M) {
  F<ST> f;
  f.s = (ST*)P.s;
  f.r = (ST*)P.r;
  f.c = nullptr;
  f.rows = P.rows;
  f.cols = P.cols;
  f.pitch = P.pitch;
  seg_fill<ST, CT>(seg, f, row, base, lane, periodic, M, RF);
}

// 1/x in FP64 for the energy weights: hardware reciprocal refined by one
// Newton step (explicit FMAs), within an ulp or two of the IEEE quotient;
// energies are reductions compared at 1e-12 relative, not bitwise.
__device__ __forceinline__ double rcp64(double x) {
  double r = __builtin_amdgcn_rcp(x);
  const double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  return r;
}

// energy term of a core point (same weights as energy_pt; core points are
// interior along SBP axes, so the H factor there is dx)
template <class CT>
__device__ __forceinline__ double energy_core(int eform, double dx, CT v, CT prm, int tx, int ty, int r, int col, int rows,
                                              int cols, int perx, int pery) {
  if (eform == 1) {
    double w = dx * dx / 2.;
    if (ty && (col == 0 || col == cols - 1)) w /= 2.;
    if (tx && (r == 0 || r == rows - 1)) w /= 2.;
    return (double)((CT)((CT)(v * prm) * v)) * w;
  }
  const double hx = (perx && tx && r == rows - 1) ? 0.0 : dx;
  const double hy = (pery && ty && col == cols - 1) ? 0.0 : dx;
  const double comp = rcp64((double)prm);
  const double w = (comp * hx) * hy;
  const double s = (double)v;
  return s * w * s;
}

// compensation variant fixed at compile time: rm, a residual array is read
// and written; km, the residual mode of value_of / add_source (kmode_of)
template <int CP>
struct CPT {
  static constexpr bool k8 = CP == TW_CPST_K8 || CP == TW_CPST_K8S;  // residual as a byte code
  static constexpr bool rm =
      CP == TW_CPST_3R || CP == TW_CPST_M3R || CP == TW_CPST_6R || CP == TW_CPST_K || k8;
  static constexpr int km = CP == TW_CPST_K ? 1 : (k8 ? 2 : 0);
  static constexpr bool rf = CP == TW_CPST_K8S;  // partner values read with their residual
};

// a lane's four residual values of row `row` at column c (multiple of 4):
// storage-type elements, or four byte codes for K8
template <class ST, int CP>
__device__ __forceinline__ V4<ST> ld_r4(const FR<ST>& X, int row, int c) {
  if constexpr (CPT<CP>::k8) {
    typedef __attribute__((address_space(1))) const uint32_t* Q;
    const uint32_t w = *(Q)(X.brow(row) + c);
    V4<ST> o;
#pragma unroll
    for (int j = 0; j < 4; j++) o.v[j] = __builtin_bit_cast(ST, (int16_t)(int8_t)(w >> (8 * j)));
    return o;
  } else {
    return ld_v4<ST>(X.rrow(row) + c);
  }
}

// ld_r4 for residual rows read once per half step (non-temporal, as ld_v4s)
template <class ST, int CP>
__device__ __forceinline__ V4<ST> ld_r4s(const FR<ST>& X, int row, int c) {
  if constexpr (CPT<CP>::k8) {
    typedef __attribute__((address_space(1))) const uint32_t* Q;
    const uint32_t w = __builtin_nontemporal_load((Q)(X.brow(row) + c));
    V4<ST> o;
#pragma unroll
    for (int j = 0; j < 4; j++) o.v[j] = __builtin_bit_cast(ST, (int16_t)(int8_t)(w >> (8 * j)));
    return o;
  } else {
    return ld_v4s<ST>(X.rrow(row) + c);
  }
}

// store of a lane's four residual values, element-wise where only some columns are in
template <class ST, int CP>
__device__ __forceinline__ void st_r4(const FR<ST>& X, int row, int c, const V4<ST>& v, bool full4, const bool* in) {
  if constexpr (CPT<CP>::k8) {
    const auto p = X.brow(row) + c;
    if (full4) {
      uint32_t w = 0;
#pragma unroll
      for (int j = 0; j < 4; j++) w |= ((uint32_t)__builtin_bit_cast(uint16_t, v.v[j]) & 0xffu) << (8 * j);
      typedef __attribute__((address_space(1))) uint32_t* Q;
      __builtin_nontemporal_store(w, (Q)p);
    } else {
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (in[j]) p[j] = (int8_t)__builtin_bit_cast(int16_t, v.v[j]);
    }
  } else {
    if (full4) {
      st_v4<ST>(X.rrow(row) + c, v);
    } else {
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (in[j]) st_1<ST>(X.rrow(row) + c + j, v.v[j]);
    }
  }
}

// accumulator start of a compensated update: R or the decoded residual, else zero
template <class ST, class CT, int CP>
__device__ __forceinline__ CT acc0(ST r, ST s) {
  if constexpr (CPT<CP>::k8)
    return (CT)k8_dec(r, s);
  else if constexpr (CPT<CP>::rm)
    return (CT)St<ST>::ld(r);
  else
    return (CT)0;
}

template <class ST, int CP>
__device__ __forceinline__ void store4(const FR<ST>& X, int r, int c, const V4<ST>& sn, const V4<ST>& rn, int full,
                                       const bool in[4]) {
  if (full) {
    st_v4<ST>(X.row(r) + c, sn);
  } else {
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (in[j]) st_1<ST>(X.row(r) + c + j, sn.v[j]);
  }
  if constexpr (CPT<CP>::rm) st_r4<ST, CP>(X, r, c, rn, full, in);
}

// Leapfrog update in registers with the compensation variant fixed at
// compile time (acc already holds R or the residual where the variant keeps
// one).  Same arithmetic as update().
template <class ST, class CT, int CP>
__device__ __forceinline__ void upd(ST& s, ST& r, CT acc) {
  using S = St<ST>;
  if constexpr (sizeof(ST) >= 4) {
    const CT a = S::ld(s);
    if constexpr (CP == TW_CPST_0) {
      s = a + acc;
    } else if constexpr (CP == TW_CPST_3R) {
      const CT t = a + acc;
      const CT z = t - a;
      r = acc - z;
      s = t;
    } else if constexpr (CP == TW_CPST_M3R) {
      const CT Y = acc;
      const CT T = a + Y;
      CT u = T - a;
      u = Y - u;
      r = u;
      s = T;
    } else {  // TW_CPST_6R
      const CT t = a + acc;
      const CT pa = t - acc;
      const CT pb = t - pa;
      const CT da = a - pa;
      const CT db = acc - pb;
      s = t;
      r = da + db;
    }
  } else {
    const CT t = S::ld(s) + acc;
    const ST hi = S::st(t);
    s = hi;
    if constexpr (CP == TW_CPST_K) r = S::st(t - S::ld(hi));
    if constexpr (CPT<CP>::k8) r = k8_enc(t - S::ld(hi), hi);
  }
}

// upd of the streaming bodies: with the 8-bit residual it also gives dq, the
// decoded value of the code it writes (q * 2^shift of the new stored value,
// exactly what k8_dec of the stored code returns), so the energy term reads
// the new value without decoding that code again.  Same results as upd.
template <class ST, class CT, int CP>
__device__ __forceinline__ void upd_q(ST& s, ST& r, CT acc, CT& dq) {
  if constexpr (sizeof(ST) == 2 && CPT<CP>::k8) {
    using S = St<ST>;
    const CT t = S::ld(s) + acc;
    const ST hi = S::st(t);
    s = hi;
    const int sh = k8_shift(hi);
    const float q = __builtin_rintf(__builtin_ldexpf(t - S::ld(hi), -sh));
    const float cq = __builtin_fminf(__builtin_fmaxf(q, -127.f), 127.f);
    r = __builtin_bit_cast(ST, (int16_t)(int)cq);
    dq = __builtin_ldexpf(cq, sh);
  } else {
    upd<ST, CT, CP>(s, r, acc);
    dq = (CT)0;
  }
}

// the new value as the energy reads it (value_of of the stored pair), from
// upd_q's dq
template <class ST, class CT, int CP>
__device__ __forceinline__ CT new_val(ST s, ST r, CT dq) {
  if constexpr (sizeof(ST) == 2 && CPT<CP>::k8)
    return (CT)St<ST>::ld(s) + dq;
  else
    return value_of<ST, CT>(s, r, CPT<CP>::km);
}

// Segment values around the lane's 4 columns: out[i] = column c - 4 + i for
// i = 2..9, the columns c - 2 .. c + 5 that dy12 reads (out[0], out[1],
// out[10], out[11] are left unset), as one 8-, one 16- and one 8-byte LDS read
template <class CT>
__device__ __forceinline__ void seg12(const CT* seg, int lane, CT out[12]) {
  const CT* p = seg + HALO + lane * VW - 4;
  if constexpr (sizeof(CT) == 4) {
    const float2 a = *(const float2*)(p + 2);
    const float4 m = *(const float4*)(p + 4);
    const float2 b = *(const float2*)(p + 8);
    out[2] = a.x;
    out[3] = a.y;
    out[4] = m.x;
    out[5] = m.y;
    out[6] = m.z;
    out[7] = m.w;
    out[8] = b.x;
    out[9] = b.y;
  } else {
#pragma unroll
    for (int q = 1; q < 5; q++) {
      const double2 v = *(const double2*)(p + 2 * q);
      out[2 * q + 0] = v.x;
      out[2 * q + 1] = v.y;
    }
  }
}

// y derivative of column j (0..3) of the lane from seg12 values
template <bool FM, class CT>
__device__ __forceinline__ CT dy12(const CT v[12], int j, int out_t, const CT st[4]) {
  const int b = 4 + j + (out_t ? -2 : -1);
  CT d = (CT)0;
#pragma unroll
  for (int k = 0; k < 4; k++) d = mac<FM>(d, v[b + k], st[k]);
  return d;
}

// source on the one row / lane that holds it (uniform row test first)
template <class ST, class CT, int CP>
__device__ __forceinline__ void src4(V4<ST>& sn, V4<ST>& rn, const Dyn& D, int grid, int r, int c) {
  if (!(D.src_on && D.src_grid == grid && r == D.src_ix)) return;
  const int j = D.src_iy - c;
#pragma unroll
  for (int q = 0; q < 4; q++) {  // selects only: no dynamic index into the register arrays
    ST s1 = sn.v[q], r1 = rn.v[q];
    add_source<ST, CT>(&s1, &r1, D.src_inc, CPT<CP>::km);
    const bool hit = q == j;
    sn.v[q] = hit ? s1 : sn.v[q];
    rn.v[q] = hit ? r1 : rn.v[q];
  }
}
// the source on one value of the source row (the caller tests the row, which
// is uniform; hit: this lane's column is the source column), keeping upd_q's
// decoded residual in step with the code
template <class ST, class CT, int CP>
__device__ __forceinline__ void src1(ST& s, ST& r, CT& dq, double inc, bool hit) {
  ST s1 = s, r1 = r;
  add_source<ST, CT>(&s1, &r1, inc, CPT<CP>::km);
  s = hit ? s1 : s;
  r = hit ? r1 : r;
  if constexpr (sizeof(ST) == 2 && CPT<CP>::k8) dq = hit ? (CT)k8_dec(r1, s1) : dq;
}

// MM energy of one cell (grid_energy.cpp:10-27, 180-201), core point (H = dx along both axes)
template <class CT>
__device__ __forceinline__ double energy_mm_core(double dx, CT lam, CT tmu, CT sxx, CT syy) {
  const double L = (double)lam, M2 = (double)tmu;
  const double rd = rcp64(M2 * (2. * L + M2));
  double p1 = (L + M2) * rd;
  double p2 = -(L * 2.) * rd;
  p1 = (p1 * dx) * dx;
  p2 = (p2 * dx) * dx;
  const double a = (double)sxx, b = (double)syy;
  return a * p1 * a + b * p1 * b + a * p2 * b;
}

// segment fill for a strip whose halo lies inside the row (no wrap, no bounds)
template <class ST, class CT, bool RF = false>
__device__ __forceinline__ void seg_fill_fast(CT* seg, const FR<ST>& P, int row, int base, int lane) {
  CT o[4];
  ld4f<ST, CT, RF>(P, row, base + lane * VW, o);
#pragma unroll
  for (int j = 0; j < 4; j++) seg[HALO + lane * VW + j] = o[j];
  if (lane < 2) {
    ld4f<ST, CT, RF>(P, row, base - HALO + lane * 4, o);
#pragma unroll
    for (int j = 0; j < 4; j++) seg[lane * 4 + j] = o[j];
  } else if (lane >= 62) {
    ld4f<ST, CT, RF>(P, row, base + STRIP + (lane - 62) * 4, o);
#pragma unroll
    for (int j = 0; j < 4; j++) seg[HALO + STRIP + (lane - 62) * 4 + j] = o[j];
  }
}

template <bool FULL, class ST, class CT, bool RF = false>
__device__ __forceinline__ void ldx_f(const FR<ST>& P, int row, int c, int periodic, int M, CT o[4]) {
  if constexpr (FULL) {
    if (periodic) row = row < 0 ? row + M : (row >= M ? row - M : row);
    if (row < 0 || row >= P.rows) {
#pragma unroll
      for (int j = 0; j < 4; j++) o[j] = (CT)0;
      return;
    }
    ld4f<ST, CT, RF>(P, row, c, o);
  } else {
    ldx_r<ST, CT, RF>(P, row, c, periodic, M, o);
  }
}

template <bool FULL, class ST, class CT, bool RF = false>
__device__ __forceinline__ void seg_f(CT* seg, const FR<ST>& P, int row, int base, int lane, int periodic, int M) {
  if constexpr (FULL)
    seg_fill_fast<ST, CT, RF>(seg, P, row, base, lane);
  else
    seg_fill_r<ST, CT, RF>(seg, P, row, base, lane, periodic, M);
}

template <bool FULL, class ST, int CP>
__device__ __forceinline__ void store_f(const FR<ST>& X, int r, int c, const V4<ST>& sn, const V4<ST>& rn, int full,
                                        const bool in[4]) {
  if constexpr (FULL) {
    st_v4<ST>(X.row(r) + c, sn);
    if constexpr (CPT<CP>::rm) st_r4<ST, CP>(X, r, c, rn, true, in);
  } else {
    store4<ST, CP>(X, r, c, sn, rn, full, in);
  }
}

// waves per SIMD the core kernels are compiled for: four (at most 128
// VGPRs) for elastic 16-bit storage and plain FP32, where latency hiding
// pays; FP32 with a residual array three; FP64 keeps its registers.  The
// acoustic 16-bit kernels issue half the loads per row, so they need more
// waves in flight: eight (at most 64 VGPRs; on the 64 batched shots of
// 512^2-2048^2, 0.726 ms per batched step against 0.747 at six and 0.768 at
// five)
template <class ST, int CP, int PH>
constexpr int WPE = sizeof(ST) == 8 ? 1
                    : sizeof(ST) == 2 && PH == TW_ACOUSTIC ? 8
                                                            : (sizeof(ST) == 4 && CP != TW_CPST_0 ? 3 : 4);

constexpr int NXCD = 8;
__device__ __forceinline__ int xcd_tile(int L, int n) {
  const int xcd = L % NXCD, local = L / NXCD;
  const int per = n / NXCD, rem = n % NXCD;
  return xcd * per + (xcd < rem ? xcd : rem) + local;
}

// ---- helpers of the FULL-strip path (strip and its halos inside the row,
// core rows only): x windows rotate through four register arrays, so the
// row loop is unrolled by four and the windows are never copied.

// row index along x with the periodic wrap (core rows never leave the grid
// otherwise)
__device__ __forceinline__ int wrap_row(int row, int periodic, int M) {
  return periodic ? (row < 0 ? row + M : (row >= M ? row - M : row)) : row;
}

template <class ST, class CT>
__device__ __forceinline__ void cvt_v4(const V4<ST>& v, CT (&o)[4]) {
#pragma unroll
  for (int j = 0; j < 4; j++) o[j] = (CT)St<ST>::ld(v.v[j]);
}

// Four coefficients of a row as loaded (codes, plus the row's affine map for
// 8-bit codes), converted later; coefficient rows are read once per half step
template <class KT>
struct Q4 {
  V4<KT> v;
  float off, step;
};
template <class KT>
__device__ __forceinline__ Q4<KT> kq4(const K<KT>& k, int row, int c) {
  Q4<KT> o;
  o.v = ld_v4s<KT>(k.p + (int64_t)row * k.pitch + c);
  if constexpr (is_code<KT>) {
    o.off = k.rp[2 * row];
    o.step = k.rp[2 * row + 1];
  }
  return o;
}
template <class KT, class CT>
__device__ __forceinline__ void kcvt4(const Q4<KT>& k, CT (&o)[4]) {
  if constexpr (is_code<KT>) {
#pragma unroll
    for (int j = 0; j < 4; j++) o[j] = c8_dec(k.v.v[j].q, k.off, k.step);
  } else {
    cvt_v4<KT, CT>(k.v, o);
  }
}
template <class KT, class CT>
__device__ __forceinline__ void kld4(const K<KT>& k, int row, int c, CT (&o)[4]) {
  kcvt4<KT, CT>(kq4<KT>(k, row, c), o);
}

// four converted values into a segment at p
template <class ST, class CT>
__device__ __forceinline__ void seg_put4(CT* p, const V4<ST>& v) {
#pragma unroll
  for (int j = 0; j < 4; j++) p[j] = (CT)St<ST>::ld(v.v[j]);
}

// core-column mask of a lane's four columns; true when all four are in
__device__ __forceinline__ bool col_mask(int c, int clo, int chi, bool (&in)[4]) {
#pragma unroll
  for (int j = 0; j < 4; j++) in[j] = c + j >= clo && c + j < chi;
  return in[0] && in[3];
}

// halo chunk column of the edge lanes (lanes 0-1 left, 62-63 right), clamped
// inside the allocation; other lanes reload their own chunk.  Clamped halo
// values only ever feed masked columns (core columns start at 4 along SBP y).
__device__ __forceinline__ int halo_col(int lane, int base, int cl, int cmax) {
  const int h = lane < 2 ? base - HALO + lane * 4 : (lane >= 62 ? base + STRIP + (lane - 62) * 4 : cl);
  return h < 0 ? 0 : (h > cmax ? cmax : h);
}

// Halo of the two LDS segments of a full-strip row, one value per lane:
// lanes 0-15 hold segment A's 16 halo values (8 columns left of the strip, 8
// right), lanes 16-31 segment B's (bodies with segment B only: lanes 0-15);
// the other lanes load a value of the row and store nothing.  Columns are
// clamped inside the row pitch (clamped values feed masked columns only).
struct Halo1 {
  int col;  // load column
  int idx;  // segment index
  int seg;  // 0: A, 1: B, -1: none
};
__device__ __forceinline__ Halo1 halo1(int lane, int base, int cmax, bool two) {
  const int h = lane & 15;
  const int c = h < 8 ? base - HALO + h : base + STRIP + h - 8;
  Halo1 o;
  o.col = c < 0 ? 0 : (c > cmax + 3 ? cmax + 3 : c);
  o.idx = h < 8 ? h : HALO + STRIP + h - 8;
  o.seg = two ? (lane < 16 ? 0 : (lane < 32 ? 1 : -1)) : (lane < 16 ? 1 : -1);
  return o;
}
// this lane's halo value: row ra of A (segment A) or rb of B, plus its
// decoded residual where the stencils read it (RF)
template <class ST, class CT, bool RF>
__device__ __forceinline__ CT halo_val(const FR<ST>& A, int ra, const FR<ST>& B, int rb, const Halo1& H) {
  const bool a = H.seg == 0;
  const ST v = ld_1<ST>((a ? A.row(ra) : B.row(rb)) + H.col);
  CT o = (CT)St<ST>::ld(v);
  if constexpr (RF && sizeof(ST) == 2) {
    const int8_t q = ld_1<int8_t>((a ? A.brow(ra) : B.brow(rb)) + H.col);
    o = o + (CT)k8_dec(__builtin_bit_cast(ST, (int16_t)q), v);
  }
  return o;
}
template <class CT>
__device__ __forceinline__ void halo_put(CT* segA, CT* segB, const Halo1& H, CT v) {
  if (H.seg == 0)
    segA[H.idx] = v;
  else if (H.seg == 1)
    segB[H.idx] = v;
}

// store of a lane's four values, element-wise where only some columns are in
template <class ST, class PT>
__device__ __forceinline__ void st_m(PT rowp, int c, const V4<ST>& v, bool full4, const bool (&in)[4]) {
  if (full4) {
    st_v4<ST>(rowp + c, v);
  } else {
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (in[j]) st_1<ST>(rowp + c + j, v.v[j]);
  }
}

// x derivative of column j from the window rows in tap order (same operation
// order as dx_win)
template <bool FM, class CT>
__device__ __forceinline__ CT dx4(const CT (&w0)[4], const CT (&w1)[4], const CT (&w2)[4], const CT (&w3)[4], int j,
                                  const CT st[4]) {
  CT v = (CT)0;
  v = mac<FM>(v, w0[j], st[0]);
  v = mac<FM>(v, w1[j], st[1]);
  v = mac<FM>(v, w2[j], st[2]);
  v = mac<FM>(v, w3[j], st[3]);
  return v;
}

// the lane's own four values of a segment row
template <class CT>
__device__ __forceinline__ void seg_own(CT* seg, const CT (&o)[4], int lane) {
#pragma unroll
  for (int j = 0; j < 4; j++) seg[HALO + lane * VW + j] = o[j];
}

// n / d in FP64 for the energy terms: hardware reciprocal, quotient and one
// residual correction (within an ulp or two of the IEEE quotient; energies are
// sums compared at 1e-12 relative)
__device__ __forceinline__ double div64(double n, double d) {
  const double r = __builtin_amdgcn_rcp(d);
  const double q = n * r;
  const double e = __builtin_fma(-d, q, n);
  return __builtin_fma(e, r, q);
}

// energy term of an interior point without its H weight dx*dx (applied to the
// block sum): s^2 / coefficient (compliance 1/coefficient for velocities with
// 1/rho, for sxy with mu and for p with kappa; grid_energy.cpp:8-216)
template <class CT>
__device__ __forceinline__ double e_single(CT v, CT prm) {
  const double s = (double)v;
  return div64(s * s, (double)prm);
}

// MM cell: [(L + M2)(a^2 + b^2) - 2 L a b] / (M2 (2 L + M2)), the quadratic
// form of the 2x2 compliance with P1 = (L+M2)/(M2(2L+M2)), P2 = -2L/(M2(2L+M2))
template <class CT>
__device__ __forceinline__ double e_mm(CT lam, CT tmu, CT sxx, CT syy) {
  const double L = (double)lam, M2 = (double)tmu;
  const double a = (double)sxx, b = (double)syy;
  const double q1 = __builtin_fma(a, a, b * b);
  const double L2 = L + L;
  const double n = __builtin_fma(L + M2, q1, -(L2 * (a * b)));
  return div64(n, M2 * (L2 + M2));
}

// Velocity half step, core region, streaming.
template <class ST, class KT, class CT, int PH, int CP>
__device__ __forceinline__ void vel3_body(const Stat<ST, KT, CT>* __restrict__ Sp, const Dyn& Dg, const Core& Cg, int bx,
                                          int by, int nbx) {
  // the step's description and the core region in registers, read before any
  // store of the launch
  const Dyn D = Dg;
  const Core C = Cg;
  constexpr bool RM = CPT<CP>::rm;
  constexpr int KM = CPT<CP>::km;
  constexpr bool RF = CPT<CP>::rf;
  constexpr bool FM = sizeof(ST) == 2;
  __shared__ __attribute__((aligned(16))) CT segA[2][SEG];  // sxy row (y partner of vx, elastic)
  __shared__ __attribute__((aligned(16))) CT segB[2][SEG];  // syy or p row (y partner of vy)
  __shared__ double sh[8];
  const Stat<ST, KT, CT>& A = *Sp;
  FR<ST> VX, VY, PX, PY, SXY;
  VX.load(A.f[TW_VX]);
  VY.load(A.f[TW_VY]);
  PX.load(A.f[PH == TW_ELASTIC ? TW_SXX : TW_P]);
  PY.load(A.f[PH == TW_ELASTIC ? TW_SYY : TW_P]);
  if constexpr (PH == TW_ELASTIC) SXY.load(A.f[TW_SXY]);
  K<KT> KBX = A.k[TW_C_BINV_VX];
  K<KT> KBY = A.k[TW_C_BINV_VY];
  {  // every array has pitch_of(cols): one pitch for the M-column grids (NM,
     // MM), one for the N-column grids (MN, NN), so row offsets are shared
    const int64_t pm = VX.pitch, pn = VY.pitch;
    PX.pitch = PY.pitch = KBX.pitch = pm;
    SXY.pitch = KBY.pitch = pn;
  }
  CT st[4];
#pragma unroll
  for (int k = 0; k < 4; k++) st[k] = A.t.st[k];
  const int perx = A.perx, pery = A.pery, mx = A.mx, my = A.my, eform = A.eform;
  const double dx = A.dx;
  const int energy = D.energy;

  const int lane = threadIdx.x;
  const int base = bx * STRIP;
  const int c = base + lane * VW;
  const int x0 = C.rlo + by * C.txc;
  const int x1 = min(x0 + C.txc, C.rhi);
  // fast path: every strip of an SBP y axis (columns masked to the core,
  // loads clamped inside the allocation), interior strips of a periodic one
  const bool inner = base >= HALO && base + STRIP + HALO <= my && base >= C.clo && base + STRIP <= C.chi;
  const bool wave_full = eform == 0 && (inner || !pery);
  const int cmax = (my + 63) / 64 * 64 - 4;  // last 4-column chunk inside every row pitch
  // block 0 finishes the previous launch's work after the solver description
  // is read: no store precedes those loads, so they stay scalar loads also when
  // the description comes from a batch table
  if (bx == 0 && by == 0) do_pending<ST, KT, CT>(A, D, sh);
  double e = 0.0;

  auto rows = [&](auto fullc) {
    constexpr bool FULL = decltype(fullc)::value;
    bool inx[4], iny[4];
    int fullx = 1, fully = 1;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int col = c + j;
      inx[j] = FULL || (col >= C.clo && col < C.chi && col < VX.cols);
      iny[j] = FULL || (col >= C.clo && col < C.chi && col < VY.cols);
      fullx &= inx[j];
      fully &= iny[j];
    }
    const bool anyx = inx[0] || inx[1] || inx[2] || inx[3];
    const bool anyy = iny[0] || iny[1] || iny[2] || iny[3];
    CT wa[4][4], wb[4][4];  // vx: x partner rows r-2..r+1 ; vy: sxy rows r-1..r+2
#pragma unroll
    for (int k = 0; k < 3; k++) ldx_f<FULL, ST, CT, RF>(PX, x0 - 2 + k, c, perx, mx, wa[k + 1]);
    if constexpr (PH == TW_ELASTIC) {
#pragma unroll
      for (int k = 0; k < 3; k++) ldx_f<FULL, ST, CT, RF>(SXY, x0 - 1 + k, c, perx, mx, wb[k + 1]);
    }
    int buf = 0;
    for (int r = x0; r < x1; r++) {
      shift_win(wa);
      ldx_f<FULL, ST, CT, RF>(PX, r + 1, c, perx, mx, wa[3]);
      if constexpr (PH == TW_ELASTIC) {
        shift_win(wb);
        ldx_f<FULL, ST, CT, RF>(SXY, r + 2, c, perx, mx, wb[3]);
        seg_f<FULL, ST, CT, RF>(segA[buf], SXY, r, base, lane, pery, my);
      }
      const bool vy_row = r < VY.rows;
      if (vy_row) seg_f<FULL, ST, CT, RF>(segB[buf], PY, r, base, lane, pery, my);
      __syncthreads();
      if (anyx && r < VX.rows) {  // ---- vx (NM)
        CT b[4];
        kld4<KT, CT>(KBX, r, c, b);
        V4<ST> sn = ld_v4<ST>(VX.row(r) + c);
        V4<ST> rn = sn;
        if constexpr (RM) rn = ld_r4<ST, CP>(VX, r, c);
        CT ya[12];
        if constexpr (PH == TW_ELASTIC) seg12<CT>(segA[buf], lane, ya);
#pragma unroll
        for (int j = 0; j < 4; j++) {
          CT acc = acc0<ST, CT, CP>(rn.v[j], sn.v[j]);
          acc = mac<FM>(acc, dx_win<FM, CT>(wa, j, st), b[j]);
          if constexpr (PH == TW_ELASTIC) acc = mac<FM>(acc, dy12<FM, CT>(ya, j, 0, st), b[j]);
          upd<ST, CT, CP>(sn.v[j], rn.v[j], acc);
        }
        src4<ST, CT, CP>(sn, rn, D, TW_GRID_NM, r, c);
        if (energy) {
#pragma unroll
          for (int j = 0; j < 4; j++) {
            const double t = energy_core<CT>(eform, dx, value_of<ST, CT>(sn.v[j], rn.v[j], KM), b[j], 1, 0, r, c + j,
                                             VX.rows, VX.cols, perx, pery);
            e += inx[j] ? t : 0.0;
          }
        }
        store_f<FULL, ST, CP>(VX, r, c, sn, rn, fullx, inx);
      }
      if (anyy && vy_row) {  // ---- vy (MN)
        CT b[4];
        kld4<KT, CT>(KBY, r, c, b);
        V4<ST> sn = ld_v4<ST>(VY.row(r) + c);
        V4<ST> rn = sn;
        if constexpr (RM) rn = ld_r4<ST, CP>(VY, r, c);
        CT yb[12];
        seg12<CT>(segB[buf], lane, yb);
#pragma unroll
        for (int j = 0; j < 4; j++) {
          CT acc = acc0<ST, CT, CP>(rn.v[j], sn.v[j]);
          if constexpr (PH == TW_ELASTIC) acc = mac<FM>(acc, dx_win<FM, CT>(wb, j, st), b[j]);
          acc = mac<FM>(acc, dy12<FM, CT>(yb, j, 1, st), b[j]);
          upd<ST, CT, CP>(sn.v[j], rn.v[j], acc);
        }
        src4<ST, CT, CP>(sn, rn, D, TW_GRID_MN, r, c);
        if (energy) {
#pragma unroll
          for (int j = 0; j < 4; j++) {
            const double t = energy_core<CT>(eform, dx, value_of<ST, CT>(sn.v[j], rn.v[j], KM), b[j], 0, 1, r, c + j,
                                             VY.rows, VY.cols, perx, pery);
            e += iny[j] ? t : 0.0;
          }
        }
        store_f<FULL, ST, CP>(VY, r, c, sn, rn, fully, iny);
      }
      buf ^= 1;
    }
  };
  // FULL strips (energy form 0).  Every load of a row is issued before any of
  // its stores, unconditionally and in bounds (halo addresses per lane,
  // clamped rows), so a row costs one memory round trip; x windows rotate,
  // segment own values come from the windows, and energy terms are summed
  // unweighted and scaled by dx*dx once.
  // F4: every column of the strip is a core column (no masks, whole-vector
  // stores, no energy selects)
  auto rows_full = [&](auto f4c) {
    constexpr bool F4 = decltype(f4c)::value;
    bool in[4] = {true, true, true, true};
    const bool full4 = F4 || col_mask(c, C.clo, C.chi, in);
    const int cl = min(c, cmax);  // load column: inside the allocation for every lane
    CT A0[4], A1[4], A2[4], A3[4];  // PX rows r-2..r+1 (x partner of vx)
    CT B0[4], B1[4], B2[4], B3[4];  // SXY rows r-1..r+2 (x partner of vy), elastic
    ldx_f<true, ST, CT, RF>(PX, x0 - 2, cl, perx, mx, A0);
    ldx_f<true, ST, CT, RF>(PX, x0 - 1, cl, perx, mx, A1);
    ldx_f<true, ST, CT, RF>(PX, x0, cl, perx, mx, A2);
    if constexpr (PH == TW_ELASTIC) {
      ldx_f<true, ST, CT, RF>(SXY, x0 - 1, cl, perx, mx, B0);
      ldx_f<true, ST, CT, RF>(SXY, x0, cl, perx, mx, B1);
      ldx_f<true, ST, CT, RF>(SXY, x0 + 1, cl, perx, mx, B2);
    }
    // segment halos: 16-bit storage one value per lane (Halo1), wider types
    // four values per edge lane (lanes 0-1 and 62-63)
    constexpr bool H1 = sizeof(ST) == 2;
    const Halo1 H = halo1(lane, base, cmax, PH == TW_ELASTIC);
    const int hc = halo_col(lane, base, cl, cmax);
    const bool hl = lane < 2 || lane >= 62;
    const int hs = lane < 2 ? lane * 4 : HALO + STRIP + (lane - 62) * 4;
    double ef = 0.0;
    auto body = [&](int r, int buf, CT(&a0)[4], CT(&a1)[4], CT(&a2)[4], CT(&a3)[4], CT(&b0)[4], CT(&b1)[4],
                    CT(&b2)[4], CT(&b3)[4]) {
      const bool vx_row = r < VX.rows, vy_row = r < VY.rows;
      const int rx = vx_row ? r : VX.rows - 1, ry = vy_row ? r : VY.rows - 1;
      // ---- loads
      const V4<ST> nA = ld_v4<ST>(PX.row(wrap_row(r + 1, perx, mx)) + cl);
      V4<ST> nB, oB;
      if constexpr (PH == TW_ELASTIC) {
        nB = ld_v4<ST>(SXY.row(wrap_row(r + 2, perx, mx)) + cl);
        oB = ld_v4<ST>(PY.row(ry) + cl);   // syy row r
      }
      // sxy row r / syy or p row r halo (acoustic: segment B only)
      CT hv = (CT)0;
      V4<ST> hA, hB;
      if constexpr (H1) {
        hv = halo_val<ST, CT, RF>(PH == TW_ELASTIC ? SXY : PY, r, PY, ry, H);
      } else {
        if constexpr (PH == TW_ELASTIC) hA = ld_v4<ST>(SXY.row(r) + hc);
        hB = ld_v4<ST>(PY.row(ry) + hc);
      }
      uint32_t cnA = 0, cnB = 0, coB = 0;  // their residual codes (K8S)
      if constexpr (RF) {
        cnA = ld_c4<ST>(PX, wrap_row(r + 1, perx, mx), cl);
        if constexpr (PH == TW_ELASTIC) {
          cnB = ld_c4<ST>(SXY, wrap_row(r + 2, perx, mx), cl);
          coB = ld_c4<ST>(PY, ry, cl);
        }
      }
      const Q4<KT> kx = kq4<KT>(KBX, rx, cl);
      V4<ST> sx = ld_v4s<ST>(VX.row(rx) + cl);
      V4<ST> qx = sx;
      if constexpr (RM) qx = ld_r4s<ST, CP>(VX, rx, cl);
      const Q4<KT> ky = kq4<KT>(KBY, ry, cl);
      V4<ST> sy = ld_v4s<ST>(VY.row(ry) + cl);
      V4<ST> qy = sy;
      if constexpr (RM) qy = ld_r4s<ST, CP>(VY, ry, cl);
      // ---- segments
      cvt_rf<ST, CT, RF>(nA, cnA, a3);
      if constexpr (PH == TW_ELASTIC) {
        cvt_rf<ST, CT, RF>(nB, cnB, b3);
        if (perx && r >= mx) {  // duplicated node row: not the window's row
          CT o[4];
          ld4f<ST, CT, RF>(SXY, r, cl, o);
          seg_own<CT>(segA[buf], o, lane);
        } else {
          seg_own<CT>(segA[buf], b1, lane);
        }
        CT o[4];
        cvt_rf<ST, CT, RF>(oB, coB, o);
        seg_own<CT>(segB[buf], o, lane);
      } else {
        seg_own<CT>(segB[buf], a2, lane);  // p row r
      }
      if constexpr (H1) {
        halo_put<CT>(segA[buf], segB[buf], H, hv);
      } else if (hl) {
        if constexpr (PH == TW_ELASTIC) seg_put4<ST, CT>(segA[buf] + hs, hA);
        seg_put4<ST, CT>(segB[buf] + hs, hB);
      }
      __syncthreads();
      if (vx_row) {  // ---- vx (NM)
        CT b[4];
        kcvt4<KT, CT>(kx, b);
        CT ya[12];
        if constexpr (PH == TW_ELASTIC) seg12<CT>(segA[buf], lane, ya);
        // per value: update, source, energy term (the duplicated node row has weight 0)
        const bool srow = D.src_on && D.src_grid == TW_GRID_NM && r == D.src_ix;
        const bool erow = energy && !(perx && r == VX.rows - 1);
#pragma unroll
        for (int j = 0; j < 4; j++) {
          CT acc = acc0<ST, CT, CP>(qx.v[j], sx.v[j]);
          acc = mac<FM>(acc, dx4<FM, CT>(a0, a1, a2, a3, j, st), b[j]);
          if constexpr (PH == TW_ELASTIC) acc = mac<FM>(acc, dy12<FM, CT>(ya, j, 0, st), b[j]);
          CT dq;
          upd_q<ST, CT, CP>(sx.v[j], qx.v[j], acc, dq);
          if (srow) src1<ST, CT, CP>(sx.v[j], qx.v[j], dq, D.src_inc, D.src_iy - c == j);
          if (erow) {
            const double t = e_single<CT>(new_val<ST, CT, CP>(sx.v[j], qx.v[j], dq), b[j]);
            ef += F4 || in[j] ? t : 0.0;
          }
        }
        st_m<ST>(VX.row(r), c, sx, full4, in);
        if constexpr (RM) st_r4<ST, CP>(VX, r, c, qx, full4, in);
      }
      if (vy_row) {  // ---- vy (MN)
        CT b[4];
        kcvt4<KT, CT>(ky, b);
        CT yb[12];
        seg12<CT>(segB[buf], lane, yb);
        const bool srow = D.src_on && D.src_grid == TW_GRID_MN && r == D.src_ix;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          CT acc = acc0<ST, CT, CP>(qy.v[j], sy.v[j]);
          if constexpr (PH == TW_ELASTIC) acc = mac<FM>(acc, dx4<FM, CT>(b0, b1, b2, b3, j, st), b[j]);
          acc = mac<FM>(acc, dy12<FM, CT>(yb, j, 1, st), b[j]);
          CT dq;
          upd_q<ST, CT, CP>(sy.v[j], qy.v[j], acc, dq);
          if (srow) src1<ST, CT, CP>(sy.v[j], qy.v[j], dq, D.src_inc, D.src_iy - c == j);
          if (energy) {
            const double t = e_single<CT>(new_val<ST, CT, CP>(sy.v[j], qy.v[j], dq), b[j]);
            ef += F4 || in[j] ? t : 0.0;
          }
        }
        st_m<ST>(VY.row(r), c, sy, full4, in);
        if constexpr (RM) st_r4<ST, CP>(VY, r, c, qy, full4, in);
      }
    };
    for (int r = x0;;) {
      if (r >= x1) break;
      body(r, 0, A0, A1, A2, A3, B0, B1, B2, B3);
      if (++r >= x1) break;
      body(r, 1, A1, A2, A3, A0, B1, B2, B3, B0);
      if (++r >= x1) break;
      body(r, 0, A2, A3, A0, A1, B2, B3, B0, B1);
      if (++r >= x1) break;
      body(r, 1, A3, A0, A1, A2, B3, B0, B1, B2);
      ++r;
    }
    e = ef * (dx * dx);
  };
  if (wave_full) {
    if (base >= C.clo && base + STRIP <= C.chi)
      rows_full(std::true_type{});
    else
      rows_full(std::false_type{});
  } else {
    rows(std::false_type{});
  }
  if (energy) {
    const double bsum = block_sum(e, sh);
    if (lane == 0) D.part[by * nbx + bx] = bsum;
  }
}

template <class ST, class KT, class CT, int PH, int CP>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE<ST, CP, PH>)))
k_vel3(const Stat<ST, KT, CT>* __restrict__ Sp, Dyn D, Core C, int nbx, int ncore, Rects R) {
  // blocks [0, nband): the band tiles of this half step, dispatched first so
  // their latency-bound per-point work overlaps the core tiles; then the
  // ncore core tiles
  const int nband = (int)gridDim.x - ncore;
  const int L = blockIdx.x;
  if (L < nband) {
    band_tile<ST, KT, CT, 0>(*Sp, D, R, L);
    return;
  }
  const int l = L - nband;
  vel3_body<ST, KT, CT, PH, CP>(Sp, D, C, l % nbx, l / nbx, nbx);
}

// Stress half step, core region, streaming.
template <class ST, class KT, class CT, int PH, int CP>
__device__ __forceinline__ void str3_body(const Stat<ST, KT, CT>* __restrict__ Sp, const Dyn& Dg, const Core& Cg, int bx,
                                          int by, int nbx) {
  // the step's description and the core region in registers, read before any
  // store of the launch
  const Dyn D = Dg;
  const Core C = Cg;
  constexpr bool RM = CPT<CP>::rm;
  constexpr int KM = CPT<CP>::km;
  constexpr bool RF = CPT<CP>::rf;
  constexpr bool FM = sizeof(ST) == 2;
  __shared__ __attribute__((aligned(16))) CT segA[2][SEG];  // vx row (y partner of sxy)
  __shared__ __attribute__((aligned(16))) CT segB[2][SEG];  // vy row (y partner of MM / p)
  __shared__ double sh[8];
  const Stat<ST, KT, CT>& A = *Sp;
  FR<ST> VX, VY, XN, X0, X1;
  VX.load(A.f[TW_VX]);
  VY.load(A.f[TW_VY]);
  X0.load(A.f[PH == TW_ELASTIC ? TW_SXX : TW_P]);
  if constexpr (PH == TW_ELASTIC) {
    X1.load(A.f[TW_SYY]);
    XN.load(A.f[TW_SXY]);
  }
  K<KT> K0 = A.k[PH == TW_ELASTIC ? TW_C_LAMBDA : TW_C_KAPPA];
  K<KT> K1 = A.k[TW_C_TWOMU];
  K<KT> KMU = A.k[TW_C_MU];
  {  // one pitch for the M-column grids (NM, MM), one for the N-column grids
     // (MN, NN): every array has pitch_of(cols)
    const int64_t pm = VX.pitch, pn = VY.pitch;
    X0.pitch = X1.pitch = K0.pitch = K1.pitch = pm;
    XN.pitch = KMU.pitch = pn;
  }
  CT st[4];
#pragma unroll
  for (int k = 0; k < 4; k++) st[k] = A.t.st[k];
  const int perx = A.perx, pery = A.pery, mx = A.mx, my = A.my, eform = A.eform;
  const double dx = A.dx;
  const int energy = D.energy;

  const int lane = threadIdx.x;
  const int base = bx * STRIP;
  const int c = base + lane * VW;
  const int x0 = C.rlo + by * C.txc;
  const int x1 = min(x0 + C.txc, C.rhi);
  // fast path: every strip of an SBP y axis (columns masked to the core,
  // loads clamped inside the allocation), interior strips of a periodic one
  const bool inner = base >= HALO && base + STRIP + HALO <= my && base >= C.clo && base + STRIP <= C.chi;
  const bool wave_full = eform == 0 && (inner || !pery);
  const int cmax = (my + 63) / 64 * 64 - 4;  // last 4-column chunk inside every row pitch
  // block 0 finishes the previous launch's work after the solver description
  // is read: no store precedes those loads, so they stay scalar loads also when
  // the description comes from a batch table
  if (bx == 0 && by == 0) do_pending<ST, KT, CT>(A, D, sh);
  double e = 0.0;

  auto rows = [&](auto fullc) {
    constexpr bool FULL = decltype(fullc)::value;
    bool inn[4], inm[4];
    int fulln = 1, fullm = 1;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int col = c + j;
      inn[j] = PH == TW_ELASTIC && (FULL || (col >= C.clo && col < C.chi && col < XN.cols));
      inm[j] = FULL || (col >= C.clo && col < C.chi && col < X0.cols);
      fulln &= inn[j];
      fullm &= inm[j];
    }
    const bool anyn = inn[0] || inn[1] || inn[2] || inn[3];
    const bool anym = inm[0] || inm[1] || inm[2] || inm[3];
    CT wa[4][4], wb[4][4];  // sxy: vy rows r-2..r+1 ; MM: vx rows r-1..r+2
    if constexpr (PH == TW_ELASTIC) {
#pragma unroll
      for (int k = 0; k < 3; k++) ldx_f<FULL, ST, CT, RF>(VY, x0 - 2 + k, c, perx, mx, wa[k + 1]);
    }
#pragma unroll
    for (int k = 0; k < 3; k++) ldx_f<FULL, ST, CT, RF>(VX, x0 - 1 + k, c, perx, mx, wb[k + 1]);
    int buf = 0;
    for (int r = x0; r < x1; r++) {
      if constexpr (PH == TW_ELASTIC) {
        shift_win(wa);
        ldx_f<FULL, ST, CT, RF>(VY, r + 1, c, perx, mx, wa[3]);
        seg_f<FULL, ST, CT, RF>(segA[buf], VX, r, base, lane, pery, my);
      }
      shift_win(wb);
      ldx_f<FULL, ST, CT, RF>(VX, r + 2, c, perx, mx, wb[3]);
      const bool mm_row = r < X0.rows;
      if (mm_row) seg_f<FULL, ST, CT, RF>(segB[buf], VY, r, base, lane, pery, my);
      __syncthreads();
      if constexpr (PH == TW_ELASTIC) {
        if (anyn && r < XN.rows) {  // ---- sxy (NN)
          CT mu[4];
          kld4<KT, CT>(KMU, r, c, mu);
          V4<ST> sn = ld_v4<ST>(XN.row(r) + c);
          V4<ST> rn = sn;
          if constexpr (RM) rn = ld_r4<ST, CP>(XN, r, c);
          CT ya[12];
          seg12<CT>(segA[buf], lane, ya);
#pragma unroll
          for (int j = 0; j < 4; j++) {
            CT acc = acc0<ST, CT, CP>(rn.v[j], sn.v[j]);
            acc = mac<FM>(acc, dx_win<FM, CT>(wa, j, st), mu[j]);
            acc = mac<FM>(acc, dy12<FM, CT>(ya, j, 1, st), mu[j]);
            upd<ST, CT, CP>(sn.v[j], rn.v[j], acc);
          }
          src4<ST, CT, CP>(sn, rn, D, TW_GRID_NN, r, c);
          if (energy) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
              const double t = energy_core<CT>(eform, dx, value_of<ST, CT>(sn.v[j], rn.v[j], KM), mu[j], 1, 1, r, c + j,
                                               XN.rows, XN.cols, perx, pery);
              e += inn[j] ? t : 0.0;
            }
          }
          store_f<FULL, ST, CP>(XN, r, c, sn, rn, fulln, inn);
        }
      }
      if (anym && mm_row) {
        CT yb[12];
        seg12<CT>(segB[buf], lane, yb);
        if constexpr (PH == TW_ELASTIC) {  // ---- sxx, syy (MM)
          CT lam[4], tmu[4];
          kld4<KT, CT>(K0, r, c, lam);
          kld4<KT, CT>(K1, r, c, tmu);
          V4<ST> n0 = ld_v4<ST>(X0.row(r) + c);
          V4<ST> n1 = ld_v4<ST>(X1.row(r) + c);
          V4<ST> m0 = n0, m1 = n1;
          if constexpr (RM) {
            m0 = ld_r4<ST, CP>(X0, r, c);
            m1 = ld_r4<ST, CP>(X1, r, c);
          }
#pragma unroll
          for (int j = 0; j < 4; j++) {
            CT a0 = acc0<ST, CT, CP>(m0.v[j], n0.v[j]);
            CT a1 = acc0<ST, CT, CP>(m1.v[j], n1.v[j]);
            const CT v = dx_win<FM, CT>(wb, j, st);
            a0 = mac<FM>(a0, v, lam[j]);
            a1 = mac<FM>(a1, v, lam[j]);
            a0 = mac<FM>(a0, v, tmu[j]);
            const CT w = dy12<FM, CT>(yb, j, 0, st);
            a0 = mac<FM>(a0, w, lam[j]);
            a1 = mac<FM>(a1, w, lam[j]);
            a1 = mac<FM>(a1, w, tmu[j]);
            upd<ST, CT, CP>(n0.v[j], m0.v[j], a0);
            upd<ST, CT, CP>(n1.v[j], m1.v[j], a1);
          }
          src4<ST, CT, CP>(n0, m0, D, TW_GRID_MM, r, c);
          src4<ST, CT, CP>(n1, m1, D, TW_GRID_MM, r, c);
          if (energy) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
              const double t = energy_mm_core<CT>(dx, lam[j], tmu[j], value_of<ST, CT>(n0.v[j], m0.v[j], KM),
                                                  value_of<ST, CT>(n1.v[j], m1.v[j], KM));
              e += inm[j] ? t : 0.0;
            }
          }
          store_f<FULL, ST, CP>(X0, r, c, n0, m0, fullm, inm);
          store_f<FULL, ST, CP>(X1, r, c, n1, m1, fullm, inm);
        } else {  // ---- p (MM, acoustic)
          CT kap[4];
          kld4<KT, CT>(K0, r, c, kap);
          V4<ST> sn = ld_v4<ST>(X0.row(r) + c);
          V4<ST> rn = sn;
          if constexpr (RM) rn = ld_r4<ST, CP>(X0, r, c);
#pragma unroll
          for (int j = 0; j < 4; j++) {
            CT acc = acc0<ST, CT, CP>(rn.v[j], sn.v[j]);
            acc = mac<FM>(acc, dx_win<FM, CT>(wb, j, st), kap[j]);
            acc = mac<FM>(acc, dy12<FM, CT>(yb, j, 0, st), kap[j]);
            upd<ST, CT, CP>(sn.v[j], rn.v[j], acc);
          }
          src4<ST, CT, CP>(sn, rn, D, TW_GRID_MM, r, c);
          if (energy) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
              const double t = energy_core<CT>(eform, dx, value_of<ST, CT>(sn.v[j], rn.v[j], KM), kap[j], 0, 0, r, c + j,
                                               X0.rows, X0.cols, perx, pery);
              e += inm[j] ? t : 0.0;
            }
          }
          store_f<FULL, ST, CP>(X0, r, c, sn, rn, fullm, inm);
        }
      }
      buf ^= 1;
    }
  };
  // FULL strips (energy form 0): as in k_vel3, all loads of a row first,
  // rotating windows, segment own values from the windows, unweighted
  // energy terms scaled by dx*dx once.
  auto rows_full = [&](auto f4c) {
    constexpr bool F4 = decltype(f4c)::value;  // every column a core column
    bool in[4] = {true, true, true, true};
    const bool full4 = F4 || col_mask(c, C.clo, C.chi, in);
    const int cl = min(c, cmax);  // load column: inside the allocation for every lane
    CT A0[4], A1[4], A2[4], A3[4];  // VY rows r-2..r+1 (x partner of sxy), elastic
    CT B0[4], B1[4], B2[4], B3[4];  // VX rows r-1..r+2 (x partner of MM)
    if constexpr (PH == TW_ELASTIC) {
      ldx_f<true, ST, CT, RF>(VY, x0 - 2, cl, perx, mx, A0);
      ldx_f<true, ST, CT, RF>(VY, x0 - 1, cl, perx, mx, A1);
      ldx_f<true, ST, CT, RF>(VY, x0, cl, perx, mx, A2);
    }
    ldx_f<true, ST, CT, RF>(VX, x0 - 1, cl, perx, mx, B0);
    ldx_f<true, ST, CT, RF>(VX, x0, cl, perx, mx, B1);
    ldx_f<true, ST, CT, RF>(VX, x0 + 1, cl, perx, mx, B2);
    // segment halos as in the velocity body
    constexpr bool H1 = sizeof(ST) == 2;
    const Halo1 H = halo1(lane, base, cmax, PH == TW_ELASTIC);
    const int hc = halo_col(lane, base, cl, cmax);
    const bool hl = lane < 2 || lane >= 62;
    const int hs = lane < 2 ? lane * 4 : HALO + STRIP + (lane - 62) * 4;
    double ef = 0.0;
    auto body = [&](int r, int buf, CT(&a0)[4], CT(&a1)[4], CT(&a2)[4], CT(&a3)[4], CT(&b0)[4], CT(&b1)[4],
                    CT(&b2)[4], CT(&b3)[4]) {
      const bool mm_row = r < X0.rows;
      const int rm = mm_row ? r : X0.rows - 1;
      // ---- loads
      V4<ST> nA, sn, qn, oB;
      Q4<KT> kn;
      if constexpr (PH == TW_ELASTIC) {
        nA = ld_v4<ST>(VY.row(wrap_row(r + 1, perx, mx)) + cl);
        kn = kq4<KT>(KMU, r, cl);
        sn = ld_v4s<ST>(XN.row(r) + cl);
        qn = sn;
        if constexpr (RM) qn = ld_r4s<ST, CP>(XN, r, cl);
      } else {
        oB = ld_v4<ST>(VY.row(rm) + cl);  // vy row r (no window in the acoustic case)
      }
      const V4<ST> nB = ld_v4<ST>(VX.row(wrap_row(r + 2, perx, mx)) + cl);
      CT hv = (CT)0;  // vx row r / vy row r halo
      V4<ST> hA, hB;
      if constexpr (H1) {
        hv = halo_val<ST, CT, RF>(VX, r, VY, rm, H);
      } else {
        if constexpr (PH == TW_ELASTIC) hA = ld_v4<ST>(VX.row(r) + hc);
        hB = ld_v4<ST>(VY.row(rm) + hc);
      }
      uint32_t cnA = 0, cnB = 0, coB = 0;  // their residual codes (K8S)
      if constexpr (RF) {
        if constexpr (PH == TW_ELASTIC)
          cnA = ld_c4<ST>(VY, wrap_row(r + 1, perx, mx), cl);
        else
          coB = ld_c4<ST>(VY, rm, cl);
        cnB = ld_c4<ST>(VX, wrap_row(r + 2, perx, mx), cl);
      }
      const Q4<KT> k0 = kq4<KT>(K0, rm, cl);
      Q4<KT> k1;
      if constexpr (PH == TW_ELASTIC) k1 = kq4<KT>(K1, rm, cl);
      V4<ST> s0 = ld_v4s<ST>(X0.row(rm) + cl), s1;
      V4<ST> q0 = s0, q1;
      if constexpr (RM) q0 = ld_r4s<ST, CP>(X0, rm, cl);
      if constexpr (PH == TW_ELASTIC) {
        s1 = ld_v4s<ST>(X1.row(rm) + cl);
        q1 = s1;
        if constexpr (RM) q1 = ld_r4s<ST, CP>(X1, rm, cl);
      }
      // ---- segments
      cvt_rf<ST, CT, RF>(nB, cnB, b3);
      if constexpr (PH == TW_ELASTIC) {
        cvt_rf<ST, CT, RF>(nA, cnA, a3);
        if (perx && r >= mx) {  // duplicated node row of vx: not the window's row
          CT o[4];
          ld4f<ST, CT, RF>(VX, r, cl, o);
          seg_own<CT>(segA[buf], o, lane);
        } else {
          seg_own<CT>(segA[buf], b1, lane);
        }
        seg_own<CT>(segB[buf], a2, lane);  // vy row r
      } else {
        CT o[4];
        cvt_rf<ST, CT, RF>(oB, coB, o);
        seg_own<CT>(segB[buf], o, lane);
      }
      if constexpr (H1) {
        halo_put<CT>(segA[buf], segB[buf], H, hv);
      } else if (hl) {
        if constexpr (PH == TW_ELASTIC) seg_put4<ST, CT>(segA[buf] + hs, hA);
        seg_put4<ST, CT>(segB[buf] + hs, hB);
      }
      __syncthreads();
      if constexpr (PH == TW_ELASTIC) {  // ---- sxy (NN); every core row is an NN row
        CT mu[4];
        kcvt4<KT, CT>(kn, mu);
        CT ya[12];
        seg12<CT>(segA[buf], lane, ya);
        // per value: update, source, energy term (the duplicated node row has weight 0)
        const bool srow = D.src_on && D.src_grid == TW_GRID_NN && r == D.src_ix;
        const bool erow = energy && !(perx && r == XN.rows - 1);
#pragma unroll
        for (int j = 0; j < 4; j++) {
          CT acc = acc0<ST, CT, CP>(qn.v[j], sn.v[j]);
          acc = mac<FM>(acc, dx4<FM, CT>(a0, a1, a2, a3, j, st), mu[j]);
          acc = mac<FM>(acc, dy12<FM, CT>(ya, j, 1, st), mu[j]);
          CT dq;
          upd_q<ST, CT, CP>(sn.v[j], qn.v[j], acc, dq);
          if (srow) src1<ST, CT, CP>(sn.v[j], qn.v[j], dq, D.src_inc, D.src_iy - c == j);
          if (erow) {
            const double t = e_single<CT>(new_val<ST, CT, CP>(sn.v[j], qn.v[j], dq), mu[j]);
            ef += F4 || in[j] ? t : 0.0;
          }
        }
        st_m<ST>(XN.row(r), c, sn, full4, in);
        if constexpr (RM) st_r4<ST, CP>(XN, r, c, qn, full4, in);
      }
      if (mm_row) {
        CT yb[12];
        seg12<CT>(segB[buf], lane, yb);
        if constexpr (PH == TW_ELASTIC) {  // ---- sxx, syy (MM)
          CT lam[4], tmu[4];
          kcvt4<KT, CT>(k0, lam);
          kcvt4<KT, CT>(k1, tmu);
          const bool srow = D.src_on && D.src_grid == TW_GRID_MM && r == D.src_ix;
#pragma unroll
          for (int j = 0; j < 4; j++) {
            CT a0 = acc0<ST, CT, CP>(q0.v[j], s0.v[j]);
            CT a1 = acc0<ST, CT, CP>(q1.v[j], s1.v[j]);
            const CT v = dx4<FM, CT>(b0, b1, b2, b3, j, st);
            a0 = mac<FM>(a0, v, lam[j]);
            a1 = mac<FM>(a1, v, lam[j]);
            a0 = mac<FM>(a0, v, tmu[j]);
            const CT w = dy12<FM, CT>(yb, j, 0, st);
            a0 = mac<FM>(a0, w, lam[j]);
            a1 = mac<FM>(a1, w, lam[j]);
            a1 = mac<FM>(a1, w, tmu[j]);
            CT d0, d1;
            upd_q<ST, CT, CP>(s0.v[j], q0.v[j], a0, d0);
            upd_q<ST, CT, CP>(s1.v[j], q1.v[j], a1, d1);
            if (srow) {
              const bool hit = D.src_iy - c == j;
              src1<ST, CT, CP>(s0.v[j], q0.v[j], d0, D.src_inc, hit);
              src1<ST, CT, CP>(s1.v[j], q1.v[j], d1, D.src_inc, hit);
            }
            if (energy) {
              const double t = e_mm<CT>(lam[j], tmu[j], new_val<ST, CT, CP>(s0.v[j], q0.v[j], d0),
                                        new_val<ST, CT, CP>(s1.v[j], q1.v[j], d1));
              ef += F4 || in[j] ? t : 0.0;
            }
          }
          st_m<ST>(X0.row(r), c, s0, full4, in);
          st_m<ST>(X1.row(r), c, s1, full4, in);
          if constexpr (RM) {
            st_r4<ST, CP>(X0, r, c, q0, full4, in);
            st_r4<ST, CP>(X1, r, c, q1, full4, in);
          }
        } else {  // ---- p (MM, acoustic)
          CT kap[4];
          kcvt4<KT, CT>(k0, kap);
          const bool srow = D.src_on && D.src_grid == TW_GRID_MM && r == D.src_ix;
#pragma unroll
          for (int j = 0; j < 4; j++) {
            CT acc = acc0<ST, CT, CP>(q0.v[j], s0.v[j]);
            acc = mac<FM>(acc, dx4<FM, CT>(b0, b1, b2, b3, j, st), kap[j]);
            acc = mac<FM>(acc, dy12<FM, CT>(yb, j, 0, st), kap[j]);
            CT dq;
            upd_q<ST, CT, CP>(s0.v[j], q0.v[j], acc, dq);
            if (srow) src1<ST, CT, CP>(s0.v[j], q0.v[j], dq, D.src_inc, D.src_iy - c == j);
            if (energy) {
              const double t = e_single<CT>(new_val<ST, CT, CP>(s0.v[j], q0.v[j], dq), kap[j]);
              ef += F4 || in[j] ? t : 0.0;
            }
          }
          st_m<ST>(X0.row(r), c, s0, full4, in);
          if constexpr (RM) st_r4<ST, CP>(X0, r, c, q0, full4, in);
        }
      }
    };
    for (int r = x0;;) {
      if (r >= x1) break;
      body(r, 0, A0, A1, A2, A3, B0, B1, B2, B3);
      if (++r >= x1) break;
      body(r, 1, A1, A2, A3, A0, B1, B2, B3, B0);
      if (++r >= x1) break;
      body(r, 0, A2, A3, A0, A1, B2, B3, B0, B1);
      if (++r >= x1) break;
      body(r, 1, A3, A0, A1, A2, B3, B0, B1, B2);
      ++r;
    }
    e = ef * (dx * dx);
  };
  if (wave_full) {
    if (base >= C.clo && base + STRIP <= C.chi)
      rows_full(std::true_type{});
    else
      rows_full(std::false_type{});
  } else {
    rows(std::false_type{});
  }
  if (energy) {
    const double bsum = block_sum(e, sh);
    if (lane == 0) D.part[by * nbx + bx] = bsum;
  }
}

template <class ST, class KT, class CT, int PH, int CP>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE<ST, CP, PH>)))
k_str3(const Stat<ST, KT, CT>* __restrict__ Sp, Dyn D, Core C, int nbx, int ncore, Rects R) {
  // blocks [0, nband): the band tiles of this half step, dispatched first so
  // their latency-bound per-point work overlaps the core tiles; then the
  // ncore core tiles
  const int nband = (int)gridDim.x - ncore;
  const int L = blockIdx.x;
  if (L < nband) {
    band_tile<ST, KT, CT, 1>(*Sp, D, R, L);
    return;
  }
  const int l = L - nband;
  str3_body<ST, KT, CT, PH, CP>(Sp, D, C, l % nbx, l / nbx, nbx);
}

// ------------------------------------------------------------ dispatch
// Core region and boundary band of the streaming launch (see Core / Rects).
static void core_geometry(const Plan& p, Core* C, Rects* R) {
  C->rlo = p.perx ? 0 : 4;
  C->rhi = p.perx ? p.mx + 1 : p.mx - 4;
  C->clo = p.pery ? 0 : 4;
  C->chi = p.pery ? p.my + 1 : p.my - 4;
  C->txc = p.txc;
  memset(R, 0, sizeof *R);
  auto add = [&](int r0, int r1, int c0, int c1) {
    const int k = R->n++;
    R->r0[k] = r0;
    R->r1[k] = r1;
    R->c0[k] = c0;
    R->c1[k] = c1;
    const int tw = tile_width(c1 - c0), th = 256 / tw;
    R->tw[k] = tw;
    R->cum[k + 1] = R->cum[k] + ((r1 - r0 + th - 1) / th) * ((c1 - c0 + tw - 1) / tw);
  };
  if (!p.perx) {
    add(0, 4, 0, p.my + 1);
    add(p.mx - 4, p.mx + 1, 0, p.my + 1);
  }
  if (!p.pery) {
    add(C->rlo, C->rhi, 0, 4);
    add(C->rlo, C->rhi, p.my - 4, p.my + 1);
  }
}

static bool use_stream(const Plan& p) {
  return p.kver != 1 && !(p.cpst == TW_CPST_3C || p.cpst == TW_CPST_M3C || p.cpst == TW_CPST_6C);
}

#if TW_PREC == 0
void grid_dims(const Plan& p, dim3* grid, dim3* block) {
  if (!use_stream(p)) {
    *block = dim3(64, 4, 1);
    *grid = dim3((unsigned)((p.my + 1 + 63) / 64), (unsigned)((p.mx + 1 + 3) / 4), 1);
    return;
  }
  Core C;
  Rects R;
  core_geometry(p, &C, &R);
  *block = dim3(64, 1, 1);
  *grid = dim3((unsigned)((C.chi + STRIP - 1) / STRIP), (unsigned)((C.rhi - C.rlo + p.txc - 1) / p.txc), 1);
}

// energy partials written per half step (core blocks + band blocks)
int half_parts(const Plan& p) {
  dim3 g, b;
  grid_dims(p, &g, &b);
  if (!use_stream(p)) return (int)(g.x * g.y);
  Core C;
  Rects R;
  core_geometry(p, &C, &R);
  return (int)(g.x * g.y) + R.cum[R.n];
}

#endif

static void grid_dims_v1(const Plan& p, dim3* grid, dim3* block) {
  *block = dim3(64, 4, 1);
  *grid = dim3((unsigned)((p.my + 1 + 63) / 64), (unsigned)((p.mx + 1 + 3) / 4), 1);
}

#if TW_PREC == 0
int energy_parts(const Plan& p) {
  dim3 g, b;
  grid_dims_v1(p, &g, &b);
  return (int)(g.x * g.y);
}
#endif

template <class ST, class KT, class CT, int PH, int CP>
static void launch_core_t(int half, const Stat<ST, KT, CT>* sp, const Dyn& d, const Core& C, dim3 g, dim3 b,
                          hipStream_t s, const Rects& R, int nband) {
  const int ncore = (int)(g.x * g.y);
  const dim3 g1((unsigned)(ncore + nband));
  if (half == 0)
    hipLaunchKernelGGL((k_vel3<ST, KT, CT, PH, CP>), g1, b, 0, s, sp, d, C, (int)g.x, ncore, R);
  else
    hipLaunchKernelGGL((k_str3<ST, KT, CT, PH, CP>), g1, b, 0, s, sp, d, C, (int)g.x, ncore, R);
}

template <class ST, class KT, class CT, int PH>
static hipError_t launch_core_p(const Plan& p, int half, const Stat<ST, KT, CT>* sp, const Dyn& d, const Core& C,
                                dim3 g, dim3 b, hipStream_t s, const Rects& R, int nband) {
  if constexpr (sizeof(ST) >= 4) {
    switch (p.cpst) {
      case TW_CPST_0: launch_core_t<ST, KT, CT, PH, TW_CPST_0>(half, sp, d, C, g, b, s, R, nband); break;
      case TW_CPST_3R: launch_core_t<ST, KT, CT, PH, TW_CPST_3R>(half, sp, d, C, g, b, s, R, nband); break;
      case TW_CPST_M3R: launch_core_t<ST, KT, CT, PH, TW_CPST_M3R>(half, sp, d, C, g, b, s, R, nband); break;
      case TW_CPST_6R: launch_core_t<ST, KT, CT, PH, TW_CPST_6R>(half, sp, d, C, g, b, s, R, nband); break;
      default: return hipErrorInvalidValue;
    }
  } else {
    if (p.cpst == TW_CPST_K)
      launch_core_t<ST, KT, CT, PH, TW_CPST_K>(half, sp, d, C, g, b, s, R, nband);
    else if (p.cpst == TW_CPST_K8)
      launch_core_t<ST, KT, CT, PH, TW_CPST_K8>(half, sp, d, C, g, b, s, R, nband);
    else if (p.cpst == TW_CPST_K8S)
      launch_core_t<ST, KT, CT, PH, TW_CPST_K8S>(half, sp, d, C, g, b, s, R, nband);
    else
      launch_core_t<ST, KT, CT, PH, TW_CPST_0>(half, sp, d, C, g, b, s, R, nband);
  }
  return hipGetLastError();
}

template <class ST, class KT, class CT>
static hipError_t launch_core(const Plan& p, int half, const Stat<ST, KT, CT>* sp, const Dyn& d, const Core& C, dim3 g,
                              dim3 b, hipStream_t s, const Rects& R, int nband) {
  if (p.physics == TW_ELASTIC) return launch_core_p<ST, KT, CT, TW_ELASTIC>(p, half, sp, d, C, g, b, s, R, nband);
  return launch_core_p<ST, KT, CT, TW_ACOUSTIC>(p, half, sp, d, C, g, b, s, R, nband);
}

template <class ST, class KT, class CT>
static hipError_t run_half(const Plan& p, int half, const StepCtl& c, hipStream_t s) {
  auto a = make_args<ST, KT, CT>(p, &c, half == 0 ? p.part_v : p.part_s);
  if (!use_stream(p)) {
    dim3 g, b;
    grid_dims_v1(p, &g, &b);
    Rects R;
    memset(&R, 0, sizeof R);
    if (half == 0)
      hipLaunchKernelGGL((k_velocity<ST, KT, CT>), g, b, 0, s, a, R);
    else
      hipLaunchKernelGGL((k_stress<ST, KT, CT>), g, b, 0, s, a, R);
    return hipGetLastError();
  }
  Core C;
  Rects R;
  core_geometry(p, &C, &R);
  dim3 g, b;
  grid_dims(p, &g, &b);
  const Stat<ST, KT, CT>* sp = (const Stat<ST, KT, CT>*)p.dstat;
  const Dyn d = a;
  // the band tiles run as extra blocks of the core launch (p.merge_bands), or
  // as a launch of their own after it
  const bool core = C.rhi > C.rlo && C.chi > C.clo;
  const bool merged = core && p.merge_bands && R.n > 0;
  if (core) {
    R.poff = (int32_t)(g.x * g.y);
    if (p.prof[2 * half]) (void)hipEventRecord(p.prof[2 * half], s);
    hipError_t e = launch_core<ST, KT, CT>(p, half, sp, d, C, g, b, s, R, merged ? R.cum[R.n] : 0);
    if (e != hipSuccess) return e;
    if (p.prof[2 * half + 1]) (void)hipEventRecord(p.prof[2 * half + 1], s);
  } else {
    // no core rows: the band launch carries the pending work
    R.poff = 0;
  }
  if (R.n > 0 && !merged) {
    R.poff = (int32_t)(g.x * g.y);
    Args<ST, KT, CT> ab = a;
    if (C.rhi > C.rlo && C.chi > C.clo) ab.pend = Pending();  // done by the core launch
    const int nb = R.cum[R.n];
    if (half == 0)
      hipLaunchKernelGGL((k_velocity<ST, KT, CT>), dim3(nb), dim3(64, 4), 0, s, ab, R);
    else
      hipLaunchKernelGGL((k_stress<ST, KT, CT>), dim3(nb), dim3(64, 4), 0, s, ab, R);
  }
  return hipGetLastError();
}

template <class ST, class KT, class CT>
static void to_other(const Plan& p, Stat<ST, KT, CT>& a) {
  for (int i = 0; i < NF; i++) {
    a.f[i].so = (ST*)p.f[i].s2;
    a.f[i].ro = (ST*)p.f[i].r2;
  }
}

template <class ST, class KT, class CT>
static size_t stat_fill_t(const Plan& p, void* out, int pp) {
  if (out) {
    Stat<ST, KT, CT> st = make_args<ST, KT, CT>(p, nullptr, nullptr);
    if (pp) to_other(p, st);
    memcpy(out, &st, sizeof st);
  }
  return sizeof(Stat<ST, KT, CT>);
}

size_t TW_PUB(stat_fill)(const Plan& p, void* out, int pp) { TW_DISPATCH(p, stat_fill_t, p, out, pp); }


template <class ST, class KT, class CT>
static hipError_t run_misc(const Plan& p, int what, const Pending* pend, hipStream_t s) {
  dim3 g, b;
  grid_dims_v1(p, &g, &b);
  StepCtl c;
  if (pend) c.pend = *pend;
  auto a = make_args<ST, KT, CT>(p, &c, p.part_s);
  if (what == 0) {
    hipLaunchKernelGGL((k_stress_energy<ST, KT, CT>), g, b, 0, s, a);
  } else if (what == 1) {
    int n = (p.mx > p.my ? p.mx : p.my) + 1;
    hipLaunchKernelGGL((k_uhat<ST, KT, CT>), dim3((n + 255) / 256), dim3(64, 4), 0, s, a);
  } else {
    hipLaunchKernelGGL((k_finalize<ST, KT, CT>), dim3(1), dim3(64, 4), 0, s, a);
  }
  return hipGetLastError();
}

hipError_t TW_PUB(launch_half)(const Plan& p, int half, const StepCtl& c, hipStream_t s) {
  TW_DISPATCH(p, run_half, p, half, c, s);
}

static hipError_t misc(const Plan& p, int what, const Pending* pend, hipStream_t s) {
  TW_DISPATCH(p, run_misc, p, what, pend, s);
}

hipError_t TW_PUB(launch_stress_energy)(const Plan& p, hipStream_t s) { return misc(p, 0, nullptr, s); }
hipError_t TW_PUB(launch_uhat)(const Plan& p, hipStream_t s) { return misc(p, 1, nullptr, s); }
hipError_t TW_PUB(launch_finalize)(const Plan& p, const Pending& pend, hipStream_t s) { return misc(p, 2, &pend, s); }

#if TW_PREC == 0
// Streaming copy (16-byte non-temporal loads and stores, grid-stride): the
// attainable read + write bandwidth of a kernel that moves every byte once,
// which the benchmark reports beside the solver's.
__global__ void __launch_bounds__(256) k_copy16(const u32x4* __restrict__ src, u32x4* __restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    typedef __attribute__((address_space(1))) const u32x4* Q;
    typedef __attribute__((address_space(1))) u32x4* W;
    const u32x4 v = __builtin_nontemporal_load((Q)(src + i));
    __builtin_nontemporal_store(v, (W)(dst + i));
  }
}

hipError_t launch_stream_copy(void* dst, const void* src, size_t bytes, hipStream_t s) {
  const size_t n = bytes / 16;
  const size_t want = (n + 255) / 256;
  const unsigned blocks = (unsigned)(want < 16384 ? (want > 0 ? want : 1) : 16384);
  hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, s, (const u32x4*)src, (u32x4*)dst, n);
  return hipGetLastError();
}

hipError_t launch_check_finite(const Plan& p, int32_t* flag, hipStream_t s) {
  dim3 b(64, 4);
  for (int f = 0; f < NF; f++) {
    const DevField& d = p.f[f];
    if (!d.s) continue;
    dim3 g((d.cols + 63) / 64, (d.rows + 3) / 4);
    switch (p.precision) {
      case TW_FP64: {
        F<double> X{(double*)d.s, nullptr, nullptr, nullptr, nullptr, d.rows, d.cols, d.pitch};
        hipLaunchKernelGGL(k_finite<double>, g, b, 0, s, X, flag);
      } break;
      case TW_FP32: {
        F<float> X{(float*)d.s, nullptr, nullptr, nullptr, nullptr, d.rows, d.cols, d.pitch};
        hipLaunchKernelGGL(k_finite<float>, g, b, 0, s, X, flag);
      } break;
      case TW_BF16: {
        F<bf16> X{(bf16*)d.s, nullptr, nullptr, nullptr, nullptr, d.rows, d.cols, d.pitch};
        hipLaunchKernelGGL(k_finite<bf16>, g, b, 0, s, X, flag);
      } break;
      default: {
        F<_Float16> X{(_Float16*)d.s, nullptr, nullptr, nullptr, nullptr, d.rows, d.cols, d.pitch};
        hipLaunchKernelGGL(k_finite<_Float16>, g, b, 0, s, X, flag);
      } break;
    }
  }
  return hipGetLastError();
}
#endif

// ------------------------------------------------------------ fused step
// One pass of the elastic leapfrog step over the interior, both axes SBP: a
// wave walks down the rows of its strip computing the velocity of row v and,
// two rows behind it, the stress of row s = v - 2 from the new velocities
// held in registers, so each field is read once and written once per step.
// Strips overlap: all 64 lanes compute velocities, lanes 1..62 (248 columns)
// own the outputs, so the stress y-derivatives of owned lanes find every new
// velocity they need inside the wave.  Chunks of rows recompute two velocity
// rows on each side.  The launch reads the old state (s pointers) and writes
// the other buffer of the ping-pong pair (so pointers); velocities of the SBP
// band rows and columns come from the band launch that ran before, in the
// new buffer.  Energies of the owned points go to part (velocity) and part2
// (stress), unweighted and scaled by dx*dx per block.
constexpr int FSTRIDE = STRIP - 2 * VW;

template <class ST, class KT, class CT, int CP>
__device__ __forceinline__ void fused_body(const Stat<ST, KT, CT>* __restrict__ Sp, const Dyn& D, const Core& C, int bx,
                                           int by, int nbx) {
  constexpr bool RM = CPT<CP>::rm;
  constexpr int KM = CPT<CP>::km;
  constexpr bool FM = sizeof(ST) == 2;
  using S = St<ST>;
  __shared__ __attribute__((aligned(16))) CT segA[SEG];  // old sxy row v (y partner of vx)
  __shared__ __attribute__((aligned(16))) CT segB[SEG];  // old syy row v (y partner of vy)
  __shared__ __attribute__((aligned(16))) CT segC[SEG];  // new vx row s (y partner of sxy)
  __shared__ __attribute__((aligned(16))) CT segD[SEG];  // new vy row s (y partner of MM)
  __shared__ double sh[8];
  const Stat<ST, KT, CT>& A = *Sp;
  FR<ST> VX, VY, SXX, SYY, SXY;
  VX.load(A.f[TW_VX]);
  VY.load(A.f[TW_VY]);
  SXX.load(A.f[TW_SXX]);
  SYY.load(A.f[TW_SYY]);
  SXY.load(A.f[TW_SXY]);
  const KT* bxp = A.k[TW_C_BINV_VX].p;
  const int64_t bxs = A.k[TW_C_BINV_VX].pitch;
  const KT* byp = A.k[TW_C_BINV_VY].p;
  const int64_t bys = A.k[TW_C_BINV_VY].pitch;
  const KT* lap = A.k[TW_C_LAMBDA].p;
  const int64_t las = A.k[TW_C_LAMBDA].pitch;
  const KT* tmp_ = A.k[TW_C_TWOMU].p;
  const int64_t tms = A.k[TW_C_TWOMU].pitch;
  const KT* mup = A.k[TW_C_MU].p;
  const int64_t mus = A.k[TW_C_MU].pitch;
  CT st[4];
#pragma unroll
  for (int k = 0; k < 4; k++) st[k] = A.t.st[k];
  const int my = A.my;
  const double dx = A.dx;
  const int energy = D.energy;

  const int lane = threadIdx.x;
  const int base = bx * FSTRIDE;
  const int c = base + lane * VW;
  const int cmax = (my + 63) / 64 * 64 - 4;
  const int cl = min(c, cmax);
  const int hc = halo_col(lane, base, cl, cmax);
  const bool hl = lane < 2 || lane >= 62;
  const int hs = lane < 2 ? lane * 4 : HALO + STRIP + (lane - 62) * 4;
  const int x0 = C.rlo + by * C.txc;
  const int x1 = min(x0 + C.txc, C.rhi);
  bool own[4], vb[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int col = c + j;
    own[j] = lane >= 1 && lane <= 62 && col >= C.clo && col < C.chi;
    vb[j] = col < C.clo || col >= C.chi;
  }
  const bool own4 = own[0] && own[3];
  const bool edge = base < C.clo || base + STRIP > C.chi;  // strip has band columns

  // x windows: sxx, new vx and syy rows as raw storage values in registers
  // (rotated by the unrolled loop); sxy and new vy rows in per-lane LDS rings
  // of five rows (each lane reads back only its own columns)
  __shared__ __attribute__((aligned(16))) V4<ST> ringB[5][64];  // old sxy rows v-2..v+2
  __shared__ __attribute__((aligned(16))) V4<ST> ringY[5][64];  // new vy rows v-4..v
  auto slot5 = [](int r) { return ((r % 5) + 5) % 5; };
  V4<ST> RA[4], RX[4], SQ[4];
  const int v0 = x0 - 2;
  RA[1] = ld_v4<ST>(SXX.row(v0 - 2) + cl);
  RA[2] = ld_v4<ST>(SXX.row(v0 - 1) + cl);
  RA[3] = ld_v4<ST>(SXX.row(v0) + cl);
  for (int r = v0 - 2; r <= v0 + 1; r++) ringB[slot5(r)][lane] = ld_v4<ST>(SXY.row(r) + cl);
  double ev = 0.0, es = 0.0;

  // roles: aF <- sxx row v+1, a0..a2 = rows v-2..v; xF <- vx row v,
  // x0..x2 = rows v-3..v-1; sqF <- syy row v, sqc = row v-2
  auto body = [&](int v, V4<ST>& aF, V4<ST>& a0, V4<ST>& a1, V4<ST>& a2, V4<ST>& xF, V4<ST>& xr0, V4<ST>& xr1,
                  V4<ST>& xr2, V4<ST>& sqF, V4<ST>& sqc) {
    const int s = v - 2;
    const bool do_s = s >= x0 && s < x1;
    const bool vcore = v >= C.rlo && v < C.rhi;
    const bool vown = v >= x0 && v < x1;
    const int sr = s < x0 ? x0 : s;  // stress row of the loads
    // ---- loads
    aF = ld_v4<ST>(SXX.row(v + 1) + cl);
    const V4<ST> nB = ld_v4<ST>(SXY.row(v + 2) + cl);
    const V4<ST> hA = ld_v4<ST>(SXY.row(v) + hc);
    sqF = ld_v4<ST>(SYY.row(v) + cl);
    const V4<ST> hB = ld_v4<ST>(SYY.row(v) + hc);
    const V4<KT> kx = ld_v4s<KT>(bxp + (int64_t)v * bxs + cl);
    V4<ST> sx = ld_v4<ST>(VX.row(v) + cl), qx = sx;
    if constexpr (RM) qx = ld_r4<ST, CP>(VX, v, cl);
    const V4<KT> ky = ld_v4s<KT>(byp + (int64_t)v * bys + cl);
    V4<ST> sy = ld_v4<ST>(VY.row(v) + cl), qy = sy;
    if constexpr (RM) qy = ld_r4<ST, CP>(VY, v, cl);
    // stress-row loads: issued with the velocity loads when the variant keeps
    // no residual (its registers allow it); otherwise after the velocity row
    V4<KT> km, k0, k1;
    V4<ST> qn, q0, q1;
    auto ld_stress = [&]() {
      km = ld_v4s<KT>(mup + (int64_t)sr * mus + cl);
      k0 = ld_v4s<KT>(lap + (int64_t)sr * las + cl);
      k1 = ld_v4s<KT>(tmp_ + (int64_t)sr * tms + cl);
      if constexpr (RM) {
        qn = ld_r4<ST, CP>(SXY, sr, cl);
        q0 = ld_r4<ST, CP>(SXX, sr, cl);
        q1 = ld_r4<ST, CP>(SYY, sr, cl);
      }
    };
    if constexpr (!RM) ld_stress();
    V4<ST> bvx, bvy;
    if (!vcore || edge) {  // the band launch's new velocities
      bvx = ld_v4<ST>(VX.orow(v) + cl);
      bvy = ld_v4<ST>(VY.orow(v) + cl);
    }
    // ---- velocity row v
    ringB[slot5(v + 2)][lane] = nB;
    {
      CT o[4];
      cvt_v4<ST, CT>(ringB[slot5(v)][lane], o);
      seg_own<CT>(segA, o, lane);
    }
    if (hl) seg_put4<ST, CT>(segA + hs, hA);
    {
      CT o[4];
      cvt_v4<ST, CT>(sqF, o);
      seg_own<CT>(segB, o, lane);
    }
    if (hl) seg_put4<ST, CT>(segB + hs, hB);
    __syncthreads();
    CT bxv[4], byv[4];
    cvt_v4<KT, CT>(kx, bxv);
    cvt_v4<KT, CT>(ky, byv);
    if (vcore) {
      CT ya[12], yb[12];
      seg12<CT>(segA, lane, ya);
      seg12<CT>(segB, lane, yb);
      CT w0[4], w1[4], w2[4], w3[4];
      cvt_v4<ST, CT>(a0, w0);
      cvt_v4<ST, CT>(a1, w1);
      cvt_v4<ST, CT>(a2, w2);
      cvt_v4<ST, CT>(aF, w3);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        CT acc = acc0<ST, CT, CP>(qx.v[j], sx.v[j]);
        acc = mac<FM>(acc, dx4<FM, CT>(w0, w1, w2, w3, j, st), bxv[j]);
        acc = mac<FM>(acc, dy12<FM, CT>(ya, j, 0, st), bxv[j]);
        upd<ST, CT, CP>(sx.v[j], qx.v[j], acc);
      }
      cvt_v4<ST, CT>(ringB[slot5(v - 1)][lane], w0);
      cvt_v4<ST, CT>(ringB[slot5(v)][lane], w1);
      cvt_v4<ST, CT>(ringB[slot5(v + 1)][lane], w2);
      cvt_v4<ST, CT>(nB, w3);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        CT acy = acc0<ST, CT, CP>(qy.v[j], sy.v[j]);
        acy = mac<FM>(acy, dx4<FM, CT>(w0, w1, w2, w3, j, st), byv[j]);
        acy = mac<FM>(acy, dy12<FM, CT>(yb, j, 1, st), byv[j]);
        upd<ST, CT, CP>(sy.v[j], qy.v[j], acy);
      }
      src4<ST, CT, CP>(sx, qx, D, TW_GRID_NM, v, c);
      src4<ST, CT, CP>(sy, qy, D, TW_GRID_MN, v, c);
      if (edge) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
          sx.v[j] = vb[j] ? bvx.v[j] : sx.v[j];
          sy.v[j] = vb[j] ? bvy.v[j] : sy.v[j];
        }
      }
    } else {
      sx = bvx;
      sy = bvy;
    }
    if (vown) {
      if (energy) {
#pragma unroll
        for (int j = 0; j < 4; j++)
          ev += own[j] ? e_single<CT>(value_of<ST, CT>(sx.v[j], qx.v[j], KM), bxv[j]) +
                             e_single<CT>(value_of<ST, CT>(sy.v[j], qy.v[j], KM), byv[j])
                       : 0.0;
      }
      st_m<ST>(VX.orow(v), c, sx, own4, own);
      st_m<ST>(VY.orow(v), c, sy, own4, own);
      if constexpr (RM) {
        st_m<ST>(VX.orrow(v), c, qx, own4, own);
        st_m<ST>(VY.orrow(v), c, qy, own4, own);
      }
    }
    xF = sx;
    // ---- stress row s = v - 2
    if (do_s) {
      if constexpr (RM) ld_stress();
      {
        CT o[4];
        cvt_v4<ST, CT>(xr1, o);
        seg_own<CT>(segC, o, lane);
        cvt_v4<ST, CT>(ringY[slot5(v - 2)][lane], o);
        seg_own<CT>(segD, o, lane);
      }
      __syncthreads();
      {  // sxy (NN)
        CT mu[4], ya[12], w0[4], w1[4], w2[4], w3[4];
        cvt_v4<KT, CT>(km, mu);
        seg12<CT>(segC, lane, ya);
        cvt_v4<ST, CT>(ringY[slot5(v - 4)][lane], w0);
        cvt_v4<ST, CT>(ringY[slot5(v - 3)][lane], w1);
        cvt_v4<ST, CT>(ringY[slot5(v - 2)][lane], w2);
        cvt_v4<ST, CT>(ringY[slot5(v - 1)][lane], w3);
        V4<ST> sn = ringB[slot5(s)][lane];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          CT acc = acc0<ST, CT, CP>(qn.v[j], sn.v[j]);
          acc = mac<FM>(acc, dx4<FM, CT>(w0, w1, w2, w3, j, st), mu[j]);
          acc = mac<FM>(acc, dy12<FM, CT>(ya, j, 1, st), mu[j]);
          upd<ST, CT, CP>(sn.v[j], qn.v[j], acc);
        }
        src4<ST, CT, CP>(sn, qn, D, TW_GRID_NN, s, c);
        if (energy) {
#pragma unroll
          for (int j = 0; j < 4; j++) es += own[j] ? e_single<CT>(value_of<ST, CT>(sn.v[j], qn.v[j], KM), mu[j]) : 0.0;
        }
        st_m<ST>(SXY.orow(s), c, sn, own4, own);
        if constexpr (RM) st_m<ST>(SXY.orrow(s), c, qn, own4, own);
      }
      {  // sxx, syy (MM)
        CT lam[4], tmu[4], yb[12], w0[4], w1[4], w2[4], w3[4];
        cvt_v4<KT, CT>(k0, lam);
        cvt_v4<KT, CT>(k1, tmu);
        seg12<CT>(segD, lane, yb);
        cvt_v4<ST, CT>(xr0, w0);
        cvt_v4<ST, CT>(xr1, w1);
        cvt_v4<ST, CT>(xr2, w2);
        cvt_v4<ST, CT>(xF, w3);
        V4<ST> s0 = a0, s1 = sqc;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          CT e0 = acc0<ST, CT, CP>(q0.v[j], s0.v[j]);
          CT e1 = acc0<ST, CT, CP>(q1.v[j], s1.v[j]);
          const CT vv = dx4<FM, CT>(w0, w1, w2, w3, j, st);
          e0 = mac<FM>(e0, vv, lam[j]);
          e1 = mac<FM>(e1, vv, lam[j]);
          e0 = mac<FM>(e0, vv, tmu[j]);
          const CT w = dy12<FM, CT>(yb, j, 0, st);
          e0 = mac<FM>(e0, w, lam[j]);
          e1 = mac<FM>(e1, w, lam[j]);
          e1 = mac<FM>(e1, w, tmu[j]);
          upd<ST, CT, CP>(s0.v[j], q0.v[j], e0);
          upd<ST, CT, CP>(s1.v[j], q1.v[j], e1);
        }
        src4<ST, CT, CP>(s0, q0, D, TW_GRID_MM, s, c);
        src4<ST, CT, CP>(s1, q1, D, TW_GRID_MM, s, c);
        if (energy) {
#pragma unroll
          for (int j = 0; j < 4; j++)
            es += own[j] ? e_mm<CT>(lam[j], tmu[j], value_of<ST, CT>(s0.v[j], q0.v[j], KM),
                                    value_of<ST, CT>(s1.v[j], q1.v[j], KM))
                         : 0.0;
        }
        st_m<ST>(SXX.orow(s), c, s0, own4, own);
        st_m<ST>(SYY.orow(s), c, s1, own4, own);
        if constexpr (RM) {
          st_m<ST>(SXX.orrow(s), c, q0, own4, own);
          st_m<ST>(SYY.orrow(s), c, q1, own4, own);
        }
      }
    }
    // the new vy row joins the ring after the stress row has read rows v-4..v-1
    ringY[slot5(v)][lane] = sy;
  };
  for (int v = v0;;) {
    if (v > x1 + 1) break;
    body(v, RA[0], RA[1], RA[2], RA[3], RX[0], RX[1], RX[2], RX[3], SQ[0], SQ[2]);
    if (++v > x1 + 1) break;
    body(v, RA[1], RA[2], RA[3], RA[0], RX[1], RX[2], RX[3], RX[0], SQ[1], SQ[3]);
    if (++v > x1 + 1) break;
    body(v, RA[2], RA[3], RA[0], RA[1], RX[2], RX[3], RX[0], RX[1], SQ[2], SQ[0]);
    if (++v > x1 + 1) break;
    body(v, RA[3], RA[0], RA[1], RA[2], RX[3], RX[0], RX[1], RX[2], SQ[3], SQ[1]);
    ++v;
  }
  if (energy) {
    const double bv = block_sum(ev * (dx * dx), sh);
    const double bs = block_sum(es * (dx * dx), sh);
    if (lane == 0) {
      D.part[by * nbx + bx] = bv;
      D.part2[by * nbx + bx] = bs;
    }
  }
}

template <class ST, class KT, class CT, int CP>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(sizeof(ST) == 8 ? 1 : (sizeof(ST) == 4 ? 2 : 3)))) k_fused(const Stat<ST, KT, CT>* __restrict__ Sp, Dyn D, Core C, int nbx) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  fused_body<ST, KT, CT, CP>(Sp, D, C, t % nbx, t / nbx, nbx);
}

// ---- fused step launches
#if TW_PREC == 0
bool fused_ok(const Plan& p) {
  return p.physics == TW_ELASTIC && !p.perx && !p.pery && p.energy_form == 0 && p.mx >= 12 && p.my >= 12 &&
         p.coef16 < 2 &&
         !(p.cpst == TW_CPST_3C || p.cpst == TW_CPST_M3C || p.cpst == TW_CPST_6C || is_k8(p.cpst));
}

#endif

static void fused_dims(const Plan& p, const Core& C, dim3* g) {
  *g = dim3((unsigned)((C.chi - C.clo + FSTRIDE - 1) / FSTRIDE), (unsigned)((C.rhi - C.rlo + p.txc - 1) / p.txc), 1);
}

#if TW_PREC == 0
int fused_parts(const Plan& p) {
  Core C;
  Rects R;
  core_geometry(p, &C, &R);
  dim3 g;
  fused_dims(p, C, &g);
  return (int)(g.x * g.y) + R.cum[R.n];
}
#endif

template <class ST, class KT, class CT>
static hipError_t run_fused(const Plan& p, int what, const StepCtl& c, hipStream_t s) {
  Core C;
  Rects R;
  core_geometry(p, &C, &R);
  dim3 g;
  fused_dims(p, C, &g);
  R.poff = (int32_t)(g.x * g.y);
  if (what == 1) {
    auto a = make_args<ST, KT, CT>(p, &c, p.part_v);
    const Dyn d = a;
    Dyn dd = d;
    dd.part2 = p.part_s;
    const Stat<ST, KT, CT>* sp = (const Stat<ST, KT, CT>*)p.dstat_pp[p.parity];
    auto go = [&](auto cp) {
      constexpr int CP = decltype(cp)::value;
      hipLaunchKernelGGL((k_fused<ST, KT, CT, CP>), dim3(g.x * g.y), dim3(64), 0, s, sp, dd, C, (int)g.x);
    };
    if constexpr (sizeof(ST) >= 4) {
      switch (p.cpst) {
        case TW_CPST_3R: go(std::integral_constant<int, TW_CPST_3R>{}); break;
        case TW_CPST_M3R: go(std::integral_constant<int, TW_CPST_M3R>{}); break;
        case TW_CPST_6R: go(std::integral_constant<int, TW_CPST_6R>{}); break;
        default: go(std::integral_constant<int, TW_CPST_0>{}); break;
      }
    } else {
      if (p.cpst == TW_CPST_K)
        go(std::integral_constant<int, TW_CPST_K>{});
      else
        go(std::integral_constant<int, TW_CPST_0>{});
    }
    return hipGetLastError();
  }
  // band launches: the velocity band reads the old state and writes the new
  // velocities; the stress band reads the new velocities and the old stress
  // and writes the new stress
  auto a = make_args<ST, KT, CT>(p, &c, what == 0 ? p.part_v : p.part_s);
  to_other(p, a);
  if (what == 2) {
    a.f[TW_VX].s = (ST*)p.f[TW_VX].s2;
    a.f[TW_VX].r = (ST*)p.f[TW_VX].r2;
    a.f[TW_VY].s = (ST*)p.f[TW_VY].s2;
    a.f[TW_VY].r = (ST*)p.f[TW_VY].r2;
  }
  if (R.n == 0) return hipSuccess;
  const int nb = R.cum[R.n];
  if (what == 0)
    hipLaunchKernelGGL((k_velocity<ST, KT, CT>), dim3(nb), dim3(64, 4), 0, s, a, R);
  else
    hipLaunchKernelGGL((k_stress<ST, KT, CT>), dim3(nb), dim3(64, 4), 0, s, a, R);
  return hipGetLastError();
}

hipError_t TW_PUB(launch_fused)(const Plan& p, int what, const StepCtl& c, hipStream_t s) {
  // the fused step takes no 8-bit coefficients (fused_ok)
  if (p.coef16 >= 2) return hipErrorInvalidValue;
#if TW_PREC == 0
  return run_fused<double, double, double>(p, what, c, s);
#elif TW_PREC == 1
  return run_fused<float, float, float>(p, what, c, s);
#elif TW_PREC == 2
  if (p.coef16) return run_fused<bf16, bf16, float>(p, what, c, s);
  return run_fused<bf16, float, float>(p, what, c, s);
#else
  if (p.coef16) return run_fused<_Float16, _Float16, float>(p, what, c, s);
  return run_fused<_Float16, float, float>(p, what, c, s);
#endif
}

// ------------------------------------------------------------ batched shots
// One launch per half step for a batch of independent problems: the blocks of
// all problems' core (or band) grids are numbered consecutively, and each
// block finds its problem by a binary search over the first-block offsets,
// then runs the same body as the single-problem kernels on that problem's
// device Stat with the step's Dyn.

// A problem's device Stat as a global-memory pointer: the table holds it as a
// plain pointer, and through a generic one the body's reads of the solver
// description would be flat (vector) loads; through a global one they are
// scalar loads, as in the single-problem kernels.
template <class ST, class KT, class CT>
__device__ __forceinline__ const Stat<ST, KT, CT>* stat_of(const void* p) {
  typedef const __attribute__((address_space(1))) Stat<ST, KT, CT>* G;
  const uint64_t u = (uint64_t)p;  // the same for every lane of the block
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
  return (const Stat<ST, KT, CT>*)(G)(((uint64_t)hi << 32) | lo);
}

// index of the problem whose block range holds `block`
__device__ __forceinline__ int batch_item(const int32_t* b0, int stride_words, int n, int block) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (b0[(int64_t)mid * stride_words] <= block)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

template <class ST, class KT, class CT, int PH, int CP>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE<ST, CP, PH>)))
k_vel3b(const BatchCore* __restrict__ items, int n, const BatchBand* __restrict__ bands, int nband,
        int core_blocks) {
  // blocks [0, band_blocks): the band tiles of all problems, dispatched first
  // (their per-point work overlaps the core tiles); then the core_blocks core
  // tiles of all problems
  const int band_blocks = (int)gridDim.x - core_blocks;
  if ((int)blockIdx.x < band_blocks) {
    const int bb = blockIdx.x;
    const BatchBand& E = bands[batch_item(&bands[0].b0, sizeof(BatchBand) / 4, nband, bb)];
    band_tile<ST, KT, CT, 0>(*stat_of<ST, KT, CT>(E.sp), E.d, E.r, bb - E.b0);
    return;
  }
  const int cb = blockIdx.x - band_blocks;
  const int k = batch_item(&items[0].b0, sizeof(BatchCore) / 4, n, cb);
  const BatchCore& B = items[k];
  const int l = cb - B.b0;
  vel3_body<ST, KT, CT, PH, CP>(stat_of<ST, KT, CT>(B.sp), B.d, B.c, l % B.nbx, l / B.nbx, B.nbx);
}

template <class ST, class KT, class CT, int PH, int CP>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE<ST, CP, PH>)))
k_str3b(const BatchCore* __restrict__ items, int n, const BatchBand* __restrict__ bands, int nband,
        int core_blocks) {
  // blocks [0, band_blocks): the band tiles of all problems, dispatched first
  // (their per-point work overlaps the core tiles); then the core_blocks core
  // tiles of all problems
  const int band_blocks = (int)gridDim.x - core_blocks;
  if ((int)blockIdx.x < band_blocks) {
    const int bb = blockIdx.x;
    const BatchBand& E = bands[batch_item(&bands[0].b0, sizeof(BatchBand) / 4, nband, bb)];
    band_tile<ST, KT, CT, 1>(*stat_of<ST, KT, CT>(E.sp), E.d, E.r, bb - E.b0);
    return;
  }
  const int cb = blockIdx.x - band_blocks;
  const int k = batch_item(&items[0].b0, sizeof(BatchCore) / 4, n, cb);
  const BatchCore& B = items[k];
  const int l = cb - B.b0;
  str3_body<ST, KT, CT, PH, CP>(stat_of<ST, KT, CT>(B.sp), B.d, B.c, l % B.nbx, l / B.nbx, B.nbx);
}

template <class ST, class KT, class CT>
__global__ void __launch_bounds__(256) k_bandb(const BatchBand* __restrict__ items, int n, int half) {
  __shared__ double sh[8];
  const int k = batch_item(&items[0].b0, sizeof(BatchBand) / 4, n, blockIdx.x);
  const BatchBand& B = items[k];
  const Stat<ST, KT, CT>& A = *stat_of<ST, KT, CT>(B.sp);
  if (half == 0)
    vel_pt<ST, KT, CT>(A, B.d, B.r, blockIdx.x - B.b0, 0, sh);
  else
    str_pt<ST, KT, CT>(A, B.d, B.r, blockIdx.x - B.b0, 0, sh);
}

template <class ST, class KT, class CT, int PH>
static void launch_batch_core(const Plan& p, int half, const BatchCore* items, int n, int blocks,
                              const BatchBand* bands, int nband, int band_blocks, hipStream_t s) {
  const dim3 g((unsigned)(blocks + band_blocks));
  auto go = [&](auto cp) {
    constexpr int CP = decltype(cp)::value;
    if (half == 0)
      hipLaunchKernelGGL((k_vel3b<ST, KT, CT, PH, CP>), g, dim3(64), 0, s, items, n, bands, nband, blocks);
    else
      hipLaunchKernelGGL((k_str3b<ST, KT, CT, PH, CP>), g, dim3(64), 0, s, items, n, bands, nband, blocks);
  };
  if constexpr (sizeof(ST) >= 4) {
    switch (p.cpst) {
      case TW_CPST_3R: go(std::integral_constant<int, TW_CPST_3R>{}); break;
      case TW_CPST_M3R: go(std::integral_constant<int, TW_CPST_M3R>{}); break;
      case TW_CPST_6R: go(std::integral_constant<int, TW_CPST_6R>{}); break;
      default: go(std::integral_constant<int, TW_CPST_0>{}); break;
    }
  } else {
    if (p.cpst == TW_CPST_K)
      go(std::integral_constant<int, TW_CPST_K>{});
    else if (p.cpst == TW_CPST_K8)
      go(std::integral_constant<int, TW_CPST_K8>{});
    else if (p.cpst == TW_CPST_K8S)
      go(std::integral_constant<int, TW_CPST_K8S>{});
    else
      go(std::integral_constant<int, TW_CPST_0>{});
  }
}

template <class ST, class KT, class CT>
static hipError_t run_batch_half(const Plan& p, int half, const BatchCore* core, int ncore, int core_blocks,
                                 const BatchBand* band, int nband, int band_blocks, hipStream_t s) {
  // the band tiles run as extra blocks of the core launch when there is one
  const bool merged = ncore > 0 && core_blocks > 0 && p.merge_bands;
  if (ncore > 0 && core_blocks > 0) {
    const int bb = merged ? band_blocks : 0;
    if (p.physics == TW_ELASTIC)
      launch_batch_core<ST, KT, CT, TW_ELASTIC>(p, half, core, ncore, core_blocks, band, nband, bb, s);
    else
      launch_batch_core<ST, KT, CT, TW_ACOUSTIC>(p, half, core, ncore, core_blocks, band, nband, bb, s);
  }
  if (nband > 0 && band_blocks > 0 && !merged)
    hipLaunchKernelGGL((k_bandb<ST, KT, CT>), dim3(band_blocks), dim3(64, 4), 0, s, band, nband, half);
  return hipGetLastError();
}

hipError_t TW_PUB(launch_batch_half)(const Plan& p, int half, const BatchCore* core, int ncore, int core_blocks,
                                     const BatchBand* band, int nband, int band_blocks, hipStream_t s) {
  TW_DISPATCH(p, run_batch_half, p, half, core, ncore, core_blocks, band, nband, band_blocks, s);
}

#if TW_PREC == 0
// batch items of one problem for one half step (host): core geometry and
// band rectangles as the single-problem launch uses them
void batch_geometry(const Plan& p, BatchCore* core, int* core_blocks, BatchBand* band, int* band_blocks) {
  Core C;
  Rects R;
  core_geometry(p, &C, &R);
  dim3 g, b;
  grid_dims(p, &g, &b);
  core->c = C;
  core->nbx = (int32_t)g.x;
  core->nb = (C.rhi > C.rlo && C.chi > C.clo) ? (int32_t)(g.x * g.y) : 0;
  *core_blocks = core->nb;
  R.poff = (int32_t)(g.x * g.y);
  band->r = R;
  band->nb = R.cum[R.n];
  *band_blocks = band->nb;
}
#endif

// ------------------------------------------------------------ adjoint
// Discrete adjoint of one acoustic leapfrog step (no compensation), the
// transpose of the linear map the forward kernels apply:
//   V: v <- v + b (G p + S(p, v))     (G: N-output derivative of p, S: SAT)
//   P: p <- p + kappa (Dv v)          (Dv: M-output derivatives of vx, vy)
// Step = P o V, so Step^T = V^T o P^T.  With mu the adjoint of the state
// after the step:
//   K1: g_kappa += mu_p (Dv v')            (v' = velocities after the step)
//       mu_v    += Dv^T (kappa mu_p)
//   K2: g_b     += mu_v (G p + S(p, v))    (p, v before the step)
//       mu_p    += G^T (b mu_v) + S_p^T (b mu_v)
//   K3: mu_v    += S_v^T (b mu_v)          (characteristic sides: -a z b)
// Gradients are accumulated in FP64; the adjoint fields are in the
// arithmetic type.

template <class CT>
struct TapT {
  const int32_t* n;
  const int32_t* idx;
  const CT* cf;
};

template <class CT>
struct AdjArgs {
  CT* mu[NF];
  double* g[6];  // by coefficient id
  double* gz[4], *gzt[4];
  CT* dn[4];
  CT* dtg[4];
  TapT<CT> gt[2], dt[2];
  const double* inj;
  int32_t nt, it_rel;
  const Receiver* rcv[2];
  int32_t nrcv[2];
  const int32_t* lead[2];  // 1 for the first receiver of its grid point in list order
  const int32_t* next[2];  // the next receiver at the same point, or -1
};

template <class ST, class KT, class CT>
__device__ __forceinline__ CT kv(const Stat<ST, KT, CT>& A, int k, int ix, int iy) {
  return kld<KT, CT>(A.k[k].get(ix, iy));
}

// A partner index along a non-periodic axis whose transposed tap list is the
// interior one: indices [5, len - 6] of a partner array of len points.  Its
// list (built on the host in ascending output order) holds the four outputs
// first + k, k = 0..3, with coefficients st[3 - k]: first = index - 2 for the
// M-output transpose (dt: an N partner read by M outputs i - 1 .. i + 2),
// index - 1 for the N-output transpose (gt: an M partner read by N outputs
// i - 2 .. i + 1).  The kernels then sum those four terms in the same order
// without reading the list.
__device__ __forceinline__ bool tap_interior(int index, int len, int periodic) {
  return !periodic && index >= 5 && index <= len - 6;
}

// acc plus the transposed taps of partner index `index` (list L, a partner
// array of len points along the axis), coefficient * f(output index) in the
// list's order; interior indices without reading the list.  FIRST is -2 for
// the M-output transposes (dt), -1 for the N-output ones (gt).
template <int FIRST, class CT, class Fn>
__device__ __forceinline__ CT tap_acc(CT acc, const TapT<CT>& L, int index, int len, int periodic, const CT st[4],
                                      Fn f) {
  if (tap_interior(index, len, periodic)) {
#pragma unroll
    for (int k = 0; k < 4; k++) acc = acc + st[3 - k] * f(index + FIRST + k);
  } else {
    for (int e = 0; e < L.n[index]; e++) acc = acc + L.cf[index * TMAX + e] * f(L.idx[index * TMAX + e]);
  }
  return acc;
}

// K1
template <class ST, class KT, class CT>
__global__ void __launch_bounds__(256) k_adj1(Args<ST, KT, CT> U, AdjArgs<CT> J) {
  const int iy = blockIdx.x * blockDim.x + threadIdx.x;
  const int ix = blockIdx.y * blockDim.y + threadIdx.y;
  const F<ST>& P = U.f[TW_P];
  const int64_t pp = P.pitch;
  if (ix < P.rows && iy < P.cols) {  // g_kappa
    const CT d = ddx<ST, CT>(U.f[TW_VX], 0, ix, iy, P.rows, U.perx, U.mx, U.t) +
                 ddy<ST, CT>(U.f[TW_VY], 0, ix, iy, P.cols, U.pery, U.my, U.t);
    J.g[TW_C_KAPPA][(int64_t)ix * pp + iy] += (double)J.mu[TW_P][(int64_t)ix * pp + iy] * (double)d;
  }
  const F<ST>& X = U.f[TW_VX];
  if (ix < X.rows && iy < X.cols) {  // mu_vx += Dx^T (kappa mu_p), along x
    CT acc = (CT)0;
    if (tap_interior(ix, X.rows, U.perx)) {
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int i = ix - 2 + k;
        acc = acc + U.t.st[3 - k] * (kv(U, TW_C_KAPPA, i, iy) * J.mu[TW_P][(int64_t)i * pp + iy]);
      }
    } else {
      const TapT<CT>& L = J.dt[0];
      for (int e = 0; e < L.n[ix]; e++) {
        const int i = L.idx[ix * TMAX + e];
        acc = acc + L.cf[ix * TMAX + e] * (kv(U, TW_C_KAPPA, i, iy) * J.mu[TW_P][(int64_t)i * pp + iy]);
      }
    }
    J.mu[TW_VX][(int64_t)ix * X.pitch + iy] += acc;
  }
  const F<ST>& Y = U.f[TW_VY];
  if (ix < Y.rows && iy < Y.cols) {  // mu_vy += Dy^T (kappa mu_p), along y
    CT acc = (CT)0;
    if (tap_interior(iy, Y.cols, U.pery)) {
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int j = iy - 2 + k;
        acc = acc + U.t.st[3 - k] * (kv(U, TW_C_KAPPA, ix, j) * J.mu[TW_P][(int64_t)ix * pp + j]);
      }
    } else {
      const TapT<CT>& L = J.dt[1];
      for (int e = 0; e < L.n[iy]; e++) {
        const int j = L.idx[iy * TMAX + e];
        acc = acc + L.cf[iy * TMAX + e] * (kv(U, TW_C_KAPPA, ix, j) * J.mu[TW_P][(int64_t)ix * pp + j]);
      }
    }
    J.mu[TW_VY][(int64_t)ix * Y.pitch + iy] += acc;
  }
}

// K2
template <class ST, class KT, class CT>
__global__ void __launch_bounds__(256) k_adj2(Args<ST, KT, CT> U, AdjArgs<CT> J) {
  using S = St<ST>;
  const int iy = blockIdx.x * blockDim.x + threadIdx.x;
  const int ix = blockIdx.y * blockDim.y + threadIdx.y;
  const F<ST>& X = U.f[TW_VX];
  const F<ST>& Y = U.f[TW_VY];
  const F<ST>& P = U.f[TW_P];
  if (ix < X.rows && iy < X.cols) {  // g_bx (and the impedance along characteristic x sides)
    const int64_t q = X.at(ix, iy);
    const CT u = (CT)S::ld(X.s[q]);
    CT t = ddx<ST, CT>(P, 1, ix, iy, X.rows, U.perx, U.mx, U.t);
    if (!U.perx) t = sat_add(U, t, TW_VX, TW_P, 0, 1, ix, iy, u, (CT)1);
    const double m = (double)J.mu[TW_VX][q];
    J.g[TW_C_BINV_VX][q] += m * (double)t;
    for (int side = 0; side < 2; side++)  // increment -a z u b: d/dz = -a u b
      if (!U.perx && U.bc[side] == 2 && ix == (side == 0 ? 0 : X.rows - 1))
        J.gz[side][iy] += m * (double)(-((side == 0 ? U.t.ainvN_L : U.t.ainvN_R) * u) * kv(U, TW_C_BINV_VX, ix, iy));
  }
  if (ix < Y.rows && iy < Y.cols) {  // g_by (and the impedance along characteristic y sides)
    const int64_t q = Y.at(ix, iy);
    const CT u = (CT)S::ld(Y.s[q]);
    CT t = ddy<ST, CT>(P, 1, ix, iy, Y.cols, U.pery, U.my, U.t);
    if (!U.pery) t = sat_add(U, t, TW_VY, TW_P, 1, 1, ix, iy, u, (CT)1);
    const double m = (double)J.mu[TW_VY][q];
    J.g[TW_C_BINV_VY][q] += m * (double)t;
    for (int side = 0; side < 2; side++)
      if (!U.pery && U.bc[2 + side] == 2 && iy == (side == 0 ? 0 : Y.cols - 1))
        J.gz[2 + side][ix] += m * (double)(-((side == 0 ? U.t.ainvN_L : U.t.ainvN_R) * u) * kv(U, TW_C_BINV_VY, ix, iy));
  }
  if (ix < P.rows && iy < P.cols) {  // mu_p += G^T (b mu_v) + SAT^T
    CT acc = (CT)0;
    if (tap_interior(ix, P.rows, U.perx)) {
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int j = ix - 1 + k;
        acc = acc + U.t.st[3 - k] * (kv(U, TW_C_BINV_VX, j, iy) * J.mu[TW_VX][X.at(j, iy)]);
      }
    } else {
      const TapT<CT>& Lx = J.gt[0];
      for (int e = 0; e < Lx.n[ix]; e++) {
        const int j = Lx.idx[ix * TMAX + e];
        acc = acc + Lx.cf[ix * TMAX + e] * (kv(U, TW_C_BINV_VX, j, iy) * J.mu[TW_VX][X.at(j, iy)]);
      }
    }
    if (tap_interior(iy, P.cols, U.pery)) {
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int j = iy - 1 + k;
        acc = acc + U.t.st[3 - k] * (kv(U, TW_C_BINV_VY, ix, j) * J.mu[TW_VY][Y.at(ix, j)]);
      }
    } else {
      const TapT<CT>& Ly = J.gt[1];
      for (int e = 0; e < Ly.n[iy]; e++) {
        const int j = Ly.idx[iy * TMAX + e];
        acc = acc + Ly.cf[iy * TMAX + e] * (kv(U, TW_C_BINV_VY, ix, j) * J.mu[TW_VY][Y.at(ix, j)]);
      }
    }
    // SAT: the boundary velocity point reads the 3-point extrapolation of p
    if (!U.perx) {
      if (U.bc[0] && ix < 3)
        acc = acc + (U.t.ainvN_L * U.t.projM_L[ix]) * (kv(U, TW_C_BINV_VX, 0, iy) * J.mu[TW_VX][X.at(0, iy)]);
      const int q = ix - (P.rows - 3);
      if (U.bc[1] && q >= 0)
        acc = acc - (U.t.ainvN_R * U.t.projM_R[q]) *
                        (kv(U, TW_C_BINV_VX, X.rows - 1, iy) * J.mu[TW_VX][X.at(X.rows - 1, iy)]);
    }
    if (!U.pery) {
      if (U.bc[2] && iy < 3)
        acc = acc + (U.t.ainvN_L * U.t.projM_L[iy]) * (kv(U, TW_C_BINV_VY, ix, 0) * J.mu[TW_VY][Y.at(ix, 0)]);
      const int q = iy - (P.cols - 3);
      if (U.bc[3] && q >= 0)
        acc = acc - (U.t.ainvN_R * U.t.projM_R[q]) *
                        (kv(U, TW_C_BINV_VY, ix, Y.cols - 1) * J.mu[TW_VY][Y.at(ix, Y.cols - 1)]);
    }
    J.mu[TW_P][P.at(ix, iy)] += acc;
  }
}

// K3: characteristic self-terms on the boundary velocity points
template <class ST, class KT, class CT>
__global__ void __launch_bounds__(256) k_adj3(Args<ST, KT, CT> U, AdjArgs<CT> J) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const F<ST>& X = U.f[TW_VX];
  const F<ST>& Y = U.f[TW_VY];
  // x sides of vx: rows 0 and mx, columns t
  if (!U.perx && t < X.cols) {
    for (int side = 0; side < 2; side++) {
      if (U.bc[side] != 2) continue;
      const int a = side == 0 ? 0 : X.rows - 1;
      const CT ai = side == 0 ? U.t.ainvN_L : U.t.ainvN_R;
      const CT c = -(ai * U.z[0][side][t]) * kv(U, TW_C_BINV_VX, a, t);
      CT& m = J.mu[TW_VX][X.at(a, t)];
      m = m + c * m;
    }
  }
  if (!U.pery && t < Y.rows) {
    for (int side = 0; side < 2; side++) {
      if (U.bc[2 + side] != 2) continue;
      const int a = side == 0 ? 0 : Y.cols - 1;
      const CT ai = side == 0 ? U.t.ainvN_L : U.t.ainvN_R;
      const CT c = -(ai * U.z[1][2 + side][t]) * kv(U, TW_C_BINV_VY, t, a);
      CT& m = J.mu[TW_VY][Y.at(t, a)];
      m = m + c * m;
    }
  }
}

struct Pitches {
  int64_t p[NF];
};

// receiver sources of forward step it_rel (relative to the adjoint run): one
// thread per receiver; the first receiver of a grid point adds those of every
// receiver at that point in list order (the chain), so the sums are those of
// a sequential loop over the lists whatever the scheduling
template <class CT>
__global__ void __launch_bounds__(256) k_adj_inject(AdjArgs<CT> J, Pitches P) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = t < J.nrcv[0] ? 0 : 1;
  const int r = g == 0 ? t : t - J.nrcv[0];
  if (r >= J.nrcv[g] || !J.lead[g][r]) return;
  const Receiver rc = J.rcv[g][r];
  CT& m = J.mu[rc.field][(int64_t)rc.ix * P.p[rc.field] + rc.iy];
  CT v = m;
  for (int q = r; q >= 0; q = J.next[g][q]) v = v + (CT)J.inj[(int64_t)J.rcv[g][q].row * J.nt + J.it_rel];
  m = v;
}

// ---- elastic
// One elastic leapfrog step (no compensation) is Step = S o V with
//   V: vx <- vx + bx (Dx_N sxx + Dy_M sxy + SAT),  vy <- vy + by (Dx_M sxy + Dy_N syy + SAT)
//   S: sxx <- sxx + (lam + 2mu) Dx_M vx + lam Dy_M vy
//      syy <- syy + lam Dx_M vx + (lam + 2mu) Dy_M vy
//      sxy <- sxy + mu (Dx_N vy + Dy_N vx)
// (D*_N: N-output derivative of M data, D*_M: M-output derivative of N data;
// SAT terms act on the velocity points only.)  A characteristic side adds to
// its normal velocity's boundary point -ai z b u (u its own value) and to each
// of its tangential velocity's 3 boundary points -ai_w z b_w uhat, uhat the
// 3-point extrapolation of those points.  Step^T = V^T o S^T:
//   E1 (velocities after the step): g_lam, g_2mu, g_mu; mu_v += D^T(C mu_s)
//   E2 (state before the step): g_bx, g_by; mu_s += G^T(b mu_v) + SAT sends^T
//   E3 (state before the step, one thread per boundary line): the
//      characteristic self-terms' increments from the incoming mu_v, and the
//      impedance sensitivities; E4 and E5 add the increments of the normal
//      and of the tangential boundary points (two launches: the corners of a
//      field are boundary points of both kinds).

// E1
template <class ST, class KT, class CT>
__global__ void __launch_bounds__(256) k_eadj1(Args<ST, KT, CT> U, AdjArgs<CT> J) {
  const int iy = blockIdx.x * blockDim.x + threadIdx.x;
  const int ix = blockIdx.y * blockDim.y + threadIdx.y;
  const F<ST>& X = U.f[TW_VX];
  const F<ST>& Y = U.f[TW_VY];
  const F<ST>& M = U.f[TW_SXX];
  const F<ST>& N = U.f[TW_SXY];
  const CT* m0 = J.mu[TW_SXX];
  const CT* m1 = J.mu[TW_SYY];
  const CT* mn = J.mu[TW_SXY];
  if (ix < M.rows && iy < M.cols) {  // g_lam, g_2mu (MM)
    const double dvx = (double)ddx<ST, CT>(X, 0, ix, iy, M.rows, U.perx, U.mx, U.t);
    const double dvy = (double)ddy<ST, CT>(Y, 0, ix, iy, M.cols, U.pery, U.my, U.t);
    const int64_t q = M.at(ix, iy);
    const double a = (double)m0[q], b = (double)m1[q];
    J.g[TW_C_LAMBDA][q] += (a + b) * (dvx + dvy);
    J.g[TW_C_TWOMU][q] += a * dvx + b * dvy;
  }
  if (ix < N.rows && iy < N.cols) {  // g_mu (NN)
    const double d = (double)ddx<ST, CT>(Y, 1, ix, iy, N.rows, U.perx, U.mx, U.t) +
                     (double)ddy<ST, CT>(X, 1, ix, iy, N.cols, U.pery, U.my, U.t);
    const int64_t q = N.at(ix, iy);
    J.g[TW_C_MU][q] += (double)mn[q] * d;
  }
  if (ix < X.rows && iy < X.cols) {  // mu_vx += Dx_M^T (lam (mu_sxx + mu_syy) + 2mu mu_sxx) + Dy_N^T (mu mu_sxy)
    CT acc = tap_acc<-2>((CT)0, J.dt[0], ix, X.rows, U.perx, U.t.st, [&](int i) {
      const int64_t q = M.at(i, iy);
      return kv(U, TW_C_LAMBDA, i, iy) * (m0[q] + m1[q]) + kv(U, TW_C_TWOMU, i, iy) * m0[q];
    });
    acc = tap_acc<-1>(acc, J.gt[1], iy, X.cols, U.pery, U.t.st,
                      [&](int j) { return kv(U, TW_C_MU, ix, j) * mn[N.at(ix, j)]; });
    J.mu[TW_VX][X.at(ix, iy)] += acc;
  }
  if (ix < Y.rows && iy < Y.cols) {  // mu_vy += Dy_M^T (lam (mu_sxx + mu_syy) + 2mu mu_syy) + Dx_N^T (mu mu_sxy)
    CT acc = tap_acc<-2>((CT)0, J.dt[1], iy, Y.cols, U.pery, U.t.st, [&](int j) {
      const int64_t q = M.at(ix, j);
      return kv(U, TW_C_LAMBDA, ix, j) * (m0[q] + m1[q]) + kv(U, TW_C_TWOMU, ix, j) * m1[q];
    });
    acc = tap_acc<-1>(acc, J.gt[0], ix, Y.rows, U.perx, U.t.st,
                      [&](int i) { return kv(U, TW_C_MU, i, iy) * mn[N.at(i, iy)]; });
    J.mu[TW_VY][Y.at(ix, iy)] += acc;
  }
}

// E2
template <class ST, class KT, class CT>
__global__ void __launch_bounds__(256) k_eadj2(Args<ST, KT, CT> U, AdjArgs<CT> J) {
  using S = St<ST>;
  const int iy = blockIdx.x * blockDim.x + threadIdx.x;
  const int ix = blockIdx.y * blockDim.y + threadIdx.y;
  const F<ST>& X = U.f[TW_VX];
  const F<ST>& Y = U.f[TW_VY];
  const F<ST>& M = U.f[TW_SXX];
  const F<ST>& N = U.f[TW_SXY];
  const CT* ux = J.mu[TW_VX];
  const CT* uy = J.mu[TW_VY];
  if (ix < X.rows && iy < X.cols) {  // g_bx: the bracket of the vx update (uhat of this state in U.uhat)
    const int64_t q = X.at(ix, iy);
    const CT u = (CT)S::ld(X.s[q]);
    CT t = ddx<ST, CT>(M, 1, ix, iy, X.rows, U.perx, U.mx, U.t);
    if (!U.perx) t = sat_add(U, t, TW_VX, TW_SXX, 0, 1, ix, iy, u, (CT)1);
    t = t + ddy<ST, CT>(N, 0, ix, iy, X.cols, U.pery, U.my, U.t);
    if (!U.pery) t = sat_add(U, t, TW_VX, TW_SXY, 1, 0, ix, iy, u, (CT)1);
    J.g[TW_C_BINV_VX][q] += (double)ux[q] * (double)t;
  }
  if (ix < Y.rows && iy < Y.cols) {  // g_by
    const int64_t q = Y.at(ix, iy);
    const CT u = (CT)S::ld(Y.s[q]);
    CT t = ddx<ST, CT>(N, 0, ix, iy, Y.rows, U.perx, U.mx, U.t);
    if (!U.perx) t = sat_add(U, t, TW_VY, TW_SXY, 0, 0, ix, iy, u, (CT)1);
    t = t + ddy<ST, CT>(U.f[TW_SYY], 1, ix, iy, Y.cols, U.pery, U.my, U.t);
    if (!U.pery) t = sat_add(U, t, TW_VY, TW_SYY, 1, 1, ix, iy, u, (CT)1);
    J.g[TW_C_BINV_VY][q] += (double)uy[q] * (double)t;
  }
  if (ix < M.rows && iy < M.cols) {
    // mu_sxx += Dx_N^T (bx mu_vx) + the sends of the vx boundary rows (3-point projection)
    CT acc = tap_acc<-1>((CT)0, J.gt[0], ix, M.rows, U.perx, U.t.st,
                         [&](int j) { return kv(U, TW_C_BINV_VX, j, iy) * ux[X.at(j, iy)]; });
    if (!U.perx) {
      if (U.bc[0] && ix < 3)
        acc = acc + (U.t.ainvN_L * U.t.projM_L[ix]) * (kv(U, TW_C_BINV_VX, 0, iy) * ux[X.at(0, iy)]);
      const int w = ix - (M.rows - 3);
      if (U.bc[1] && w >= 0)
        acc = acc - (U.t.ainvN_R * U.t.projM_R[w]) *
                        (kv(U, TW_C_BINV_VX, X.rows - 1, iy) * ux[X.at(X.rows - 1, iy)]);
    }
    J.mu[TW_SXX][M.at(ix, iy)] += acc;
    // mu_syy += Dy_N^T (by mu_vy) + the sends of the vy boundary columns
    acc = tap_acc<-1>((CT)0, J.gt[1], iy, M.cols, U.pery, U.t.st,
                      [&](int j) { return kv(U, TW_C_BINV_VY, ix, j) * uy[Y.at(ix, j)]; });
    if (!U.pery) {
      if (U.bc[2] && iy < 3)
        acc = acc + (U.t.ainvN_L * U.t.projM_L[iy]) * (kv(U, TW_C_BINV_VY, ix, 0) * uy[Y.at(ix, 0)]);
      const int w = iy - (M.cols - 3);
      if (U.bc[3] && w >= 0)
        acc = acc - (U.t.ainvN_R * U.t.projM_R[w]) *
                        (kv(U, TW_C_BINV_VY, ix, Y.cols - 1) * uy[Y.at(ix, Y.cols - 1)]);
    }
    J.mu[TW_SYY][M.at(ix, iy)] += acc;
  }
  if (ix < N.rows && iy < N.cols) {
    // mu_sxy += Dy_M^T (bx mu_vx) + Dx_M^T (by mu_vy) + the sends read by the
    // tangential velocities' 3 boundary points of each side
    CT acc = tap_acc<-2>((CT)0, J.dt[1], iy, N.cols, U.pery, U.t.st,
                         [&](int j) { return kv(U, TW_C_BINV_VX, ix, j) * ux[X.at(ix, j)]; });
    acc = tap_acc<-2>(acc, J.dt[0], ix, N.rows, U.perx, U.t.st,
                      [&](int i) { return kv(U, TW_C_BINV_VY, i, iy) * uy[Y.at(i, iy)]; });
    if (!U.pery && ix < X.rows) {
      if (U.bc[2] && iy == 0)
        for (int w = 0; w < 3; w++) acc = acc + U.t.ainvM_L[w] * (kv(U, TW_C_BINV_VX, ix, w) * ux[X.at(ix, w)]);
      if (U.bc[3] && iy == N.cols - 1)
        for (int w = 0; w < 3; w++) {
          const int c = X.cols - 3 + w;
          acc = acc - U.t.ainvM_R[w] * (kv(U, TW_C_BINV_VX, ix, c) * ux[X.at(ix, c)]);
        }
    }
    if (!U.perx && iy < Y.cols) {
      if (U.bc[0] && ix == 0)
        for (int w = 0; w < 3; w++) acc = acc + U.t.ainvM_L[w] * (kv(U, TW_C_BINV_VY, w, iy) * uy[Y.at(w, iy)]);
      if (U.bc[1] && ix == N.rows - 1)
        for (int w = 0; w < 3; w++) {
          const int r = Y.rows - 3 + w;
          acc = acc - U.t.ainvM_R[w] * (kv(U, TW_C_BINV_VY, r, iy) * uy[Y.at(r, iy)]);
        }
    }
    J.mu[TW_SXY][N.at(ix, iy)] += acc;
  }
}

// E3: per characteristic side and boundary line t, from the incoming adjoint
template <class ST, class KT, class CT>
__global__ void __launch_bounds__(256) k_eadj3(Args<ST, KT, CT> U, AdjArgs<CT> J) {
  using S = St<ST>;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  for (int side = 0; side < 4; side++) {
    if (U.bc[side] != 2) continue;
    const int axis = side >> 1, right = side & 1;
    {  // normal velocity: vx on the x sides, vy on the y sides (N-type along the axis)
      const int f = axis == 0 ? TW_VX : TW_VY;
      const F<ST>& V = U.f[f];
      if (t < (axis == 0 ? V.cols : V.rows)) {
        const int a = right ? (axis == 0 ? V.rows : V.cols) - 1 : 0;
        const int ix = axis == 0 ? a : t, iy = axis == 0 ? t : a;
        const int64_t q = V.at(ix, iy);
        const CT ai = right ? U.t.ainvN_R : U.t.ainvN_L;
        const CT b = kv(U, f == TW_VX ? TW_C_BINV_VX : TW_C_BINV_VY, ix, iy);
        const CT m = J.mu[f][q];
        const CT u = (CT)S::ld(V.s[q]);
        J.dn[side][t] = (-(ai * U.z[f == TW_VX ? 0 : 1][side][t]) * b) * m;
        J.gz[side][t] += (double)m * (double)(-(ai * u) * b);
      }
    }
    {  // tangential velocity: vy on the x sides, vx on the y sides (M-type along the axis)
      const int f = axis == 0 ? TW_VY : TW_VX;
      const F<ST>& V = U.f[f];
      const int n_along = axis == 0 ? V.rows : V.cols;
      if (t < (axis == 0 ? V.cols : V.rows)) {
        const CT* pm = right ? U.t.projM_R : U.t.projM_L;
        const CT* aiw = right ? U.t.ainvM_R : U.t.ainvM_L;
        const CT z = U.z[f == TW_VX ? 0 : 1][side][t];
        CT uh = (CT)0;  // as k_uhat
        for (int w = 0; w < 3; w++) {
          const int a = right ? n_along - 3 + w : w;
          uh = uh + pm[w] * (CT)S::ld(V.s[axis == 0 ? V.at(a, t) : V.at(t, a)]);
        }
        CT sum = (CT)0;
        double gs = 0.0;
        for (int w = 0; w < 3; w++) {
          const int a = right ? n_along - 3 + w : w;
          const int ix = axis == 0 ? a : t, iy = axis == 0 ? t : a;
          const CT b = kv(U, f == TW_VX ? TW_C_BINV_VX : TW_C_BINV_VY, ix, iy);
          const CT m = J.mu[f][V.at(ix, iy)];
          sum = sum + (-(aiw[w] * z) * b) * m;
          gs += (double)m * (double)(-(aiw[w] * uh) * b);
        }
        for (int w = 0; w < 3; w++) J.dtg[side][3 * t + w] = pm[w] * sum;
        J.gzt[side][t] += gs;
      }
    }
  }
}

// E4: the normal velocities' boundary points; E5 (tangential = 1): the
// tangential velocities' 3 boundary points of each line
template <class ST, class KT, class CT>
__global__ void __launch_bounds__(256) k_eadj45(Args<ST, KT, CT> U, AdjArgs<CT> J, int tangential) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  for (int side = 0; side < 4; side++) {
    if (U.bc[side] != 2) continue;
    const int axis = side >> 1, right = side & 1;
    const int f = (axis == 0) == !tangential ? TW_VX : TW_VY;
    const F<ST>& V = U.f[f];
    if (t >= (axis == 0 ? V.cols : V.rows)) continue;
    const int n_along = axis == 0 ? V.rows : V.cols;
    if (!tangential) {
      const int a = right ? n_along - 1 : 0;
      CT& m = J.mu[f][axis == 0 ? V.at(a, t) : V.at(t, a)];
      m = m + J.dn[side][t];
    } else {
      for (int w = 0; w < 3; w++) {
        const int a = right ? n_along - 3 + w : w;
        CT& m = J.mu[f][axis == 0 ? V.at(a, t) : V.at(t, a)];
        m = m + J.dtg[side][3 * t + w];
      }
    }
  }
}

template <class ST, class KT, class CT>
static hipError_t run_adjoint_step(const Plan& p, const AdjPlan& a, const StateView& cur, const StateView& prev,
                                   int it, int it_rel, hipStream_t s) {
  AdjArgs<CT> J;
  for (int f = 0; f < NF; f++) J.mu[f] = (CT*)a.mu[f];
  for (int i = 0; i < 6; i++) J.g[i] = a.g[i];
  for (int i = 0; i < 4; i++) {
    J.gz[i] = a.gz[i];
    J.gzt[i] = a.gzt[i];
    J.dn[i] = (CT*)a.dn[i];
    J.dtg[i] = (CT*)a.dtg[i];
  }
  for (int d = 0; d < 2; d++) {
    J.gt[d] = TapT<CT>{a.gt[d].n, a.gt[d].idx, (const CT*)a.gt[d].cf};
    J.dt[d] = TapT<CT>{a.dt[d].n, a.dt[d].idx, (const CT*)a.dt[d].cf};
  }
  J.inj = a.inj;
  J.nt = a.nt;
  J.it_rel = it_rel;
  J.rcv[0] = p.rcv_v;
  J.rcv[1] = p.rcv_s;
  J.nrcv[0] = p.nrcv_v;
  J.nrcv[1] = p.nrcv_s;
  for (int g = 0; g < 2; g++) {
    J.lead[g] = a.lead[g];
    J.next[g] = a.next[g];
  }
  StepCtl c;
  c.it = it;
  auto U1 = make_args<ST, KT, CT>(p, &c, nullptr);
  auto U0 = U1;
  for (int f = 0; f < NF; f++) {
    if (cur.s[f]) U1.f[f].s = (ST*)cur.s[f];
    if (prev.s[f]) U0.f[f].s = (ST*)prev.s[f];
  }
  Pitches P;
  for (int f = 0; f < NF; f++) P.p[f] = p.f[f].pitch;
  const int nr = p.nrcv_v + p.nrcv_s;
  if (nr > 0) hipLaunchKernelGGL((k_adj_inject<CT>), dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, s, J, P);
  const dim3 b(64, 4), g((unsigned)((p.my + 1 + 63) / 64), (unsigned)((p.mx + 1 + 3) / 4));
  const int n = (p.mx > p.my ? p.mx : p.my) + 1;
  const dim3 gl((unsigned)((n + 255) / 256));
  if (p.physics == TW_ACOUSTIC) {
    hipLaunchKernelGGL((k_adj1<ST, KT, CT>), g, b, 0, s, U1, J);
    hipLaunchKernelGGL((k_adj2<ST, KT, CT>), g, b, 0, s, U0, J);
    hipLaunchKernelGGL((k_adj3<ST, KT, CT>), gl, dim3(256), 0, s, U0, J);
    return hipGetLastError();
  }
  const bool chr = p.bc[0] == 'C' || p.bc[1] == 'C' || p.bc[2] == 'C' || p.bc[3] == 'C';
  hipLaunchKernelGGL((k_eadj1<ST, KT, CT>), g, b, 0, s, U1, J);
  // the tangential velocities' boundary extrapolation of the state before the
  // step, which the SAT terms of E2 read (the forward step recomputes it)
  if (chr) hipLaunchKernelGGL((k_uhat<ST, KT, CT>), gl, dim3(64, 4), 0, s, U0);
  hipLaunchKernelGGL((k_eadj2<ST, KT, CT>), g, b, 0, s, U0, J);
  if (chr) {
    hipLaunchKernelGGL((k_eadj3<ST, KT, CT>), gl, dim3(256), 0, s, U0, J);
    hipLaunchKernelGGL((k_eadj45<ST, KT, CT>), gl, dim3(256), 0, s, U0, J, 0);
    hipLaunchKernelGGL((k_eadj45<ST, KT, CT>), gl, dim3(256), 0, s, U0, J, 1);
  }
  return hipGetLastError();
}

hipError_t TW_PUB(launch_adjoint_step)(const Plan& p, const AdjPlan& a, const StateView& cur, const StateView& prev,
                                       int it, int it_rel, hipStream_t s) {
  // the adjoint runs FP64 and FP32 storage only
#if TW_PREC <= 1
  TW_DISPATCH(p, run_adjoint_step, p, a, cur, prev, it, it_rel, s);
#else
  (void)p, (void)a, (void)cur, (void)prev, (void)it, (void)it_rel, (void)s;
  return hipErrorInvalidValue;
#endif
}

}  // namespace tw
