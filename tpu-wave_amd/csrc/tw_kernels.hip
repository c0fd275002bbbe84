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

#include "tw_internal.h"

namespace tw {

// ------------------------------------------------------------ storage types
struct bf16 {
  uint16_t u;
};

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
  static __device__ __forceinline__ bf16 st(float f) {
    uint32_t u = __float_as_uint(f);
    bf16 r;
    if ((u & 0x7fffffffu) > 0x7f800000u) {
      r.u = (uint16_t)((u >> 16) | 0x40u);
      return r;
    }
    u += 0x7fffu + ((u >> 16) & 1u);
    r.u = (uint16_t)(u >> 16);
    return r;
  }
};
template <>
struct St<_Float16> {
  using C = float;
  static __device__ __forceinline__ float ld(_Float16 v) { return (float)v; }
  static __device__ __forceinline__ _Float16 st(float f) { return (_Float16)f; }
};

template <class KT, class CT>
__device__ __forceinline__ CT kld(KT v) {
  return (CT)St<KT>::ld(v);
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
  ST* s;
  ST* r;
  ST* c;
  int32_t rows, cols;
  int64_t pitch;
  __device__ __forceinline__ int64_t at(int ix, int iy) const { return (int64_t)ix * pitch + iy; }
};

template <class KT>
struct K {
  const KT* p;
  int64_t pitch;
  __device__ __forceinline__ KT get(int ix, int iy) const { return p[(int64_t)ix * pitch + iy]; }
};

template <class ST, class KT, class CT>
struct Args {
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
  // step control
  int32_t it, energy;
  int32_t src_on, src_grid, src_ix, src_iy;
  double src_inc;
  double* part;
  Pending pend;
  double* rec;
  int32_t nt_max;
};

template <class ST, class KT, class CT>
static Args<ST, KT, CT> make_args(const Plan& p, const StepCtl* c, double* part) {
  Args<ST, KT, CT> a;
  for (int i = 0; i < NF; i++) {
    a.f[i].s = (ST*)p.f[i].s;
    a.f[i].r = (ST*)p.f[i].r;
    a.f[i].c = (ST*)p.f[i].c;
    a.f[i].rows = p.f[i].rows;
    a.f[i].cols = p.f[i].cols;
    a.f[i].pitch = p.f[i].pitch;
    a.k[i].p = (const KT*)p.k[i].p;
    a.k[i].pitch = p.k[i].pitch;
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
// Derivative along x onto output row ix of an output with n rows, type out_t
// (1 = N, 0 = M), from partner P.  v = 0; v = v + S*c per tap in tap order.
template <class ST, class CT>
__device__ __forceinline__ CT ddx(const F<ST>& P, int out_t, int ix, int iy, int n, int periodic, int M,
                                  const TabT<CT>& T) {
  using S = St<ST>;
  CT v = (CT)0;
  if (periodic) {
    int r = ix + (out_t ? -2 : -1);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      int rr = r + k;
      rr = rr < 0 ? rr + M : (rr >= M ? rr - M : rr);
      v = v + (CT)S::ld(P.s[P.at(rr, iy)]) * T.st[k];
    }
    return v;
  }
  const int nb = out_t ? 4 : 3;
  if (ix < nb) {
#pragma unroll
    for (int w = 0; w < 5; w++) v = v + (CT)S::ld(P.s[P.at(w, iy)]) * (out_t ? T.DN_L[ix][w] : T.DM_L[ix][w]);
  } else if (ix >= n - nb) {
    const int r = ix - (n - nb);
    const int b = P.rows - 5;
#pragma unroll
    for (int w = 0; w < 5; w++) v = v + (CT)S::ld(P.s[P.at(b + w, iy)]) * (out_t ? T.DN_R[r][w] : T.DM_R[r][w]);
  } else {
    int r = ix + (out_t ? -2 : -1);
#pragma unroll
    for (int k = 0; k < 4; k++) v = v + (CT)S::ld(P.s[P.at(r + k, iy)]) * T.st[k];
  }
  return v;
}

template <class ST, class CT>
__device__ __forceinline__ CT ddy(const F<ST>& P, int out_t, int ix, int iy, int n, int periodic, int M,
                                  const TabT<CT>& T) {
  using S = St<ST>;
  const ST* row = P.s + (int64_t)ix * P.pitch;
  CT v = (CT)0;
  if (periodic) {
    int c = iy + (out_t ? -2 : -1);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      int cc = c + k;
      cc = cc < 0 ? cc + M : (cc >= M ? cc - M : cc);
      v = v + (CT)S::ld(row[cc]) * T.st[k];
    }
    return v;
  }
  const int nb = out_t ? 4 : 3;
  if (iy < nb) {
#pragma unroll
    for (int w = 0; w < 5; w++) v = v + (CT)S::ld(row[w]) * (out_t ? T.DN_L[iy][w] : T.DM_L[iy][w]);
  } else if (iy >= n - nb) {
    const int r = iy - (n - nb);
    const int b = P.cols - 5;
#pragma unroll
    for (int w = 0; w < 5; w++) v = v + (CT)S::ld(row[b + w]) * (out_t ? T.DN_R[r][w] : T.DM_R[r][w]);
  } else {
    int c = iy + (out_t ? -2 : -1);
#pragma unroll
    for (int k = 0; k < 4; k++) v = v + (CT)S::ld(row[c + k]) * T.st[k];
  }
  return v;
}

// SAT term on a velocity point (secure_bdry, cuda/src/cpp/grid_drvt.tpp:469-695).
// axis 0: boundary rows, j = column; axis 1: boundary columns, j = row.
// own: the point's old value (u_hat for an N-type output); uh: precomputed
// extrapolation for an M-type output.  Returns the increment to add to acc,
// already multiplied by the coefficient, or leaves acc unchanged.
template <class ST, class KT, class CT>
__device__ __forceinline__ CT sat_add(const Args<ST, KT, CT>& A, CT acc, int vf, int pf, int axis, int out_t, int ix,
                                      int iy, CT own, CT coef) {
  using S = St<ST>;
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
      const CT sv = axis == 0 ? (CT)S::ld(P.s[P.at(pa, j)]) : (CT)S::ld(P.s[P.at(j, pa)]);
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
  return (CT)0;
}

// Compensated update (cuda/src/cuda/grid_update.cut:62-124).  Writes the state
// (and companion) and returns the value the point now represents.
template <class ST, class CT>
__device__ __forceinline__ void update(const F<ST>& X, int64_t q, CT acc, int cpst, ST* snew, ST* rnew) {
  using S = St<ST>;
  if constexpr (sizeof(ST) >= 4) {
    const CT s = S::ld(X.s[q]);
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
    const CT t = S::ld(X.s[q]) + acc;
    const ST hi = S::st(t);
    *snew = hi;
    if (cpst == TW_CPST_K) *rnew = S::st(t - S::ld(hi));
  }
}

// source: S = (T)((double)S + inc) after the update (forward_simulation.cpp:57-114)
template <class ST, class CT>
__device__ __forceinline__ void add_source(ST* snew, ST* rnew, double inc, int kmode) {
  using S = St<ST>;
  if constexpr (sizeof(ST) >= 4) {
    *snew = (ST)((double)*snew + inc);
  } else {
    if (kmode) {
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
    if (kmode) v = v + (CT)St<ST>::ld(r);
  }
  return v;
}

// H-norm factor along an axis (adjust_parameters_energy_periodic,
// cuda/src/cpp/grid_energy.cpp:58-177)
template <class ST, class KT, class CT>
__device__ __forceinline__ double hw(const Args<ST, KT, CT>& A, int axis, int t, int i, int n) {
  const int periodic = axis == 0 ? A.perx : A.pery;
  if (periodic) return (t && i == n - 1) ? 0.0 : A.dx;
  const int nb = t ? 4 : 3;
  if (i < nb) return t ? A.hN[i] : A.hM[i];
  if (i >= n - nb) return t ? A.hN[n - 1 - i] : A.hM[n - 1 - i];
  return A.dx;
}

// energy term of one scalar field point (grid_energy.cpp:180-216 / cpu_acoustic.cpp:327-344)
template <class ST, class KT, class CT>
__device__ __forceinline__ double energy_pt(const Args<ST, KT, CT>& A, CT v, CT prm, int tx, int ty, int ix, int iy,
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
__device__ void do_pending(const Args<ST, KT, CT>& A, double* sh) {
  const Pending& P = A.pend;
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
  const int kmode = A.cpst == TW_CPST_K;
  for (int r = tid; r < P.nrcv; r += nth) {
    const Receiver rc = P.rcv[r];
    const F<ST>& X = A.f[rc.field];
    const int64_t q = X.at(rc.ix, rc.iy);
    ST rr = X.s[q];
    if (kmode) rr = X.r[q];
    A.rec[(int64_t)rc.row * A.nt_max + P.rcv_it] = (double)value_of<ST, CT>(X.s[q], rr, kmode);
  }
}

// ------------------------------------------------------------ half steps
// Velocity half step: vx (NM) and vy (MN).
template <class ST, class KT, class CT>
__global__ void __launch_bounds__(256) k_velocity(Args<ST, KT, CT> A) {
  __shared__ double sh[8];
  if (blockIdx.x == 0 && blockIdx.y == 0) do_pending(A, sh);
  const int iy = blockIdx.x * blockDim.x + threadIdx.x;
  const int ix = blockIdx.y * blockDim.y + threadIdx.y;
  const int el = A.physics == TW_ELASTIC;
  const int kmode = A.cpst == TW_CPST_K;
  double e = 0.0;
  // ---- vx on NM: d/dx of sxx (elastic) or p, + d/dy of sxy (elastic)
  {
    const F<ST>& X = A.f[TW_VX];
    if (ix < X.rows && iy < X.cols) {
      const int64_t q = X.at(ix, iy);
      const CT b = kld<KT, CT>(A.k[TW_C_BINV_VX].get(ix, iy));
      const CT own = (CT)St<ST>::ld(X.s[q]);
      CT acc = acc_init<ST, CT>(X, q, A.cpst);
      const int pxf = el ? TW_SXX : TW_P;
      CT d = ddx<ST, CT>(A.f[pxf], 1, ix, iy, X.rows, A.perx, A.mx, A.t);
      acc = acc + d * b;
      if (!A.perx) acc = sat_add(A, acc, TW_VX, pxf, 0, 1, ix, iy, own, b);
      if (el) {
        d = ddy<ST, CT>(A.f[TW_SXY], 0, ix, iy, X.cols, A.pery, A.my, A.t);
        acc = acc + d * b;
        if (!A.pery) acc = sat_add(A, acc, TW_VX, TW_SXY, 1, 0, ix, iy, own, b);
      }
      ST sn = X.s[q], rn = sn;
      if (kmode || A.cpst == TW_CPST_3R || A.cpst == TW_CPST_M3R || A.cpst == TW_CPST_6R) rn = X.r[q];
      update<ST, CT>(X, q, acc, A.cpst, &sn, &rn);
      if (A.src_on && A.src_grid == TW_GRID_NM && ix == A.src_ix && iy == A.src_iy)
        add_source<ST, CT>(&sn, &rn, A.src_inc, kmode);
      X.s[q] = sn;
      if (kmode || A.cpst == TW_CPST_3R || A.cpst == TW_CPST_M3R || A.cpst == TW_CPST_6R) X.r[q] = rn;
      if (A.energy) e += energy_pt(A, value_of<ST, CT>(sn, rn, kmode), b, 1, 0, ix, iy, X.rows, X.cols);
    }
  }
  // ---- vy on MN: d/dx of sxy (elastic) + d/dy of syy (elastic) or p
  {
    const F<ST>& X = A.f[TW_VY];
    if (ix < X.rows && iy < X.cols) {
      const int64_t q = X.at(ix, iy);
      const CT b = kld<KT, CT>(A.k[TW_C_BINV_VY].get(ix, iy));
      const CT own = (CT)St<ST>::ld(X.s[q]);
      CT acc = acc_init<ST, CT>(X, q, A.cpst);
      if (el) {
        CT d = ddx<ST, CT>(A.f[TW_SXY], 0, ix, iy, X.rows, A.perx, A.mx, A.t);
        acc = acc + d * b;
        if (!A.perx) acc = sat_add(A, acc, TW_VY, TW_SXY, 0, 0, ix, iy, own, b);
      }
      const int pyf = el ? TW_SYY : TW_P;
      CT d = ddy<ST, CT>(A.f[pyf], 1, ix, iy, X.cols, A.pery, A.my, A.t);
      acc = acc + d * b;
      if (!A.pery) acc = sat_add(A, acc, TW_VY, pyf, 1, 1, ix, iy, own, b);
      ST sn = X.s[q], rn = sn;
      if (kmode || A.cpst == TW_CPST_3R || A.cpst == TW_CPST_M3R || A.cpst == TW_CPST_6R) rn = X.r[q];
      update<ST, CT>(X, q, acc, A.cpst, &sn, &rn);
      if (A.src_on && A.src_grid == TW_GRID_MN && ix == A.src_ix && iy == A.src_iy)
        add_source<ST, CT>(&sn, &rn, A.src_inc, kmode);
      X.s[q] = sn;
      if (kmode || A.cpst == TW_CPST_3R || A.cpst == TW_CPST_M3R || A.cpst == TW_CPST_6R) X.r[q] = rn;
      if (A.energy) e += energy_pt(A, value_of<ST, CT>(sn, rn, kmode), b, 0, 1, ix, iy, X.rows, X.cols);
    }
  }
  if (A.energy) {
    const double bsum = block_sum(e, sh);
    if (threadIdx.x == 0 && threadIdx.y == 0) A.part[blockIdx.y * gridDim.x + blockIdx.x] = bsum;
  }
}

// Stress half step: sxy (NN) and sxx, syy (MM), or p (MM, acoustic).
template <class ST, class KT, class CT>
__global__ void __launch_bounds__(256) k_stress(Args<ST, KT, CT> A) {
  __shared__ double sh[8];
  if (blockIdx.x == 0 && blockIdx.y == 0) do_pending(A, sh);
  const int iy = blockIdx.x * blockDim.x + threadIdx.x;
  const int ix = blockIdx.y * blockDim.y + threadIdx.y;
  const int kmode = A.cpst == TW_CPST_K;
  const int rmode = kmode || A.cpst == TW_CPST_3R || A.cpst == TW_CPST_M3R || A.cpst == TW_CPST_6R;
  double e = 0.0;
  if (A.physics == TW_ELASTIC) {
    {
      const F<ST>& X = A.f[TW_SXY];
      if (ix < X.rows && iy < X.cols) {
        const int64_t q = X.at(ix, iy);
        const CT mu = kld<KT, CT>(A.k[TW_C_MU].get(ix, iy));
        CT acc = acc_init<ST, CT>(X, q, A.cpst);
        CT d = ddx<ST, CT>(A.f[TW_VY], 1, ix, iy, X.rows, A.perx, A.mx, A.t);
        acc = acc + d * mu;
        d = ddy<ST, CT>(A.f[TW_VX], 1, ix, iy, X.cols, A.pery, A.my, A.t);
        acc = acc + d * mu;
        ST sn = X.s[q], rn = sn;
        if (rmode) rn = X.r[q];
        update<ST, CT>(X, q, acc, A.cpst, &sn, &rn);
        if (A.src_on && A.src_grid == TW_GRID_NN && ix == A.src_ix && iy == A.src_iy)
          add_source<ST, CT>(&sn, &rn, A.src_inc, kmode);
        X.s[q] = sn;
        if (rmode) X.r[q] = rn;
        if (A.energy) e += energy_pt(A, value_of<ST, CT>(sn, rn, kmode), mu, 1, 1, ix, iy, X.rows, X.cols);
      }
    }
    {
      const F<ST>& X0 = A.f[TW_SXX];
      const F<ST>& X1 = A.f[TW_SYY];
      if (ix < X0.rows && iy < X0.cols) {
        const int64_t q = X0.at(ix, iy);
        const CT lam = kld<KT, CT>(A.k[TW_C_LAMBDA].get(ix, iy));
        const CT tmu = kld<KT, CT>(A.k[TW_C_TWOMU].get(ix, iy));
        CT a0 = acc_init<ST, CT>(X0, q, A.cpst);
        CT a1 = acc_init<ST, CT>(X1, q, A.cpst);
        CT v = ddx<ST, CT>(A.f[TW_VX], 0, ix, iy, X0.rows, A.perx, A.mx, A.t);
        a0 = a0 + v * lam;
        a1 = a1 + v * lam;
        a0 = a0 + v * tmu;
        CT w = ddy<ST, CT>(A.f[TW_VY], 0, ix, iy, X0.cols, A.pery, A.my, A.t);
        a0 = a0 + w * lam;
        a1 = a1 + w * lam;
        a1 = a1 + w * tmu;
        ST s0 = X0.s[q], r0 = s0, s1 = X1.s[q], r1 = s1;
        if (rmode) {
          r0 = X0.r[q];
          r1 = X1.r[q];
        }
        update<ST, CT>(X0, q, a0, A.cpst, &s0, &r0);
        update<ST, CT>(X1, q, a1, A.cpst, &s1, &r1);
        if (A.src_on && A.src_grid == TW_GRID_MM && ix == A.src_ix && iy == A.src_iy) {
          add_source<ST, CT>(&s0, &r0, A.src_inc, kmode);
          add_source<ST, CT>(&s1, &r1, A.src_inc, kmode);
        }
        X0.s[q] = s0;
        X1.s[q] = s1;
        if (rmode) {
          X0.r[q] = r0;
          X1.r[q] = r1;
        }
        if (A.energy) {
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
    if (ix < X.rows && iy < X.cols) {
      const int64_t q = X.at(ix, iy);
      const CT kap = kld<KT, CT>(A.k[TW_C_KAPPA].get(ix, iy));
      CT acc = acc_init<ST, CT>(X, q, A.cpst);
      CT d = ddx<ST, CT>(A.f[TW_VX], 0, ix, iy, X.rows, A.perx, A.mx, A.t);
      acc = acc + d * kap;
      d = ddy<ST, CT>(A.f[TW_VY], 0, ix, iy, X.cols, A.pery, A.my, A.t);
      acc = acc + d * kap;
      ST sn = X.s[q], rn = sn;
      if (rmode) rn = X.r[q];
      update<ST, CT>(X, q, acc, A.cpst, &sn, &rn);
      if (A.src_on && A.src_grid == TW_GRID_MM && ix == A.src_ix && iy == A.src_iy)
        add_source<ST, CT>(&sn, &rn, A.src_inc, kmode);
      X.s[q] = sn;
      if (rmode) X.r[q] = rn;
      if (A.energy) {
        const CT v = value_of<ST, CT>(sn, rn, kmode);
        if (A.eform == 1) e += energy_pt(A, v, kap, 0, 0, ix, iy, X.rows, X.cols);
        else e += energy_pt(A, v, kap, 0, 0, ix, iy, X.rows, X.cols);
      }
    }
  }
  if (A.energy) {
    const double bsum = block_sum(e, sh);
    if (threadIdx.x == 0 && threadIdx.y == 0) A.part[blockIdx.y * gridDim.x + blockIdx.x] = bsum;
  }
}

// Energy of the stress fields only (start of a run / after the state was set).
template <class ST, class KT, class CT>
__global__ void __launch_bounds__(256) k_stress_energy(Args<ST, KT, CT> A) {
  __shared__ double sh[8];
  const int iy = blockIdx.x * blockDim.x + threadIdx.x;
  const int ix = blockIdx.y * blockDim.y + threadIdx.y;
  const int kmode = A.cpst == TW_CPST_K;
  double e = 0.0;
  if (A.physics == TW_ELASTIC) {
    const F<ST>& X = A.f[TW_SXY];
    if (ix < X.rows && iy < X.cols) {
      const int64_t q = X.at(ix, iy);
      const CT mu = kld<KT, CT>(A.k[TW_C_MU].get(ix, iy));
      ST r = X.s[q];
      if (kmode) r = X.r[q];
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
        r0 = X0.r[q];
        r1 = X1.r[q];
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
      if (kmode) r = X.r[q];
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
  using S = St<ST>;
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
      const CT v = axis == 0 ? (CT)S::ld(X.s[X.at(a, j)]) : (CT)S::ld(X.s[X.at(j, a)]);
      u = u + pm[w] * v;
    }
    out[j] = u;
  }
}

template <class ST, class KT, class CT>
__global__ void __launch_bounds__(256) k_finalize(Args<ST, KT, CT> A) {
  __shared__ double sh[8];
  do_pending(A, sh);
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

// ------------------------------------------------------------ dispatch
void grid_dims(const Plan& p, dim3* grid, dim3* block) {
  *block = dim3(64, 4, 1);
  *grid = dim3((unsigned)((p.my + 1 + 63) / 64), (unsigned)((p.mx + 1 + 3) / 4), 1);
}

template <class ST, class KT, class CT>
static hipError_t run_half(const Plan& p, int half, const StepCtl& c, hipStream_t s) {
  dim3 g, b;
  grid_dims(p, &g, &b);
  auto a = make_args<ST, KT, CT>(p, &c, half == 0 ? p.part_v : p.part_s);
  if (half == 0)
    hipLaunchKernelGGL((k_velocity<ST, KT, CT>), g, b, 0, s, a);
  else
    hipLaunchKernelGGL((k_stress<ST, KT, CT>), g, b, 0, s, a);
  return hipGetLastError();
}

template <class ST, class KT, class CT>
static hipError_t run_misc(const Plan& p, int what, const Pending* pend, hipStream_t s) {
  dim3 g, b;
  grid_dims(p, &g, &b);
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

hipError_t launch_half(const Plan& p, int half, const StepCtl& c, hipStream_t s) {
  switch (p.precision) {
    case TW_FP64: return run_half<double, double, double>(p, half, c, s);
    case TW_FP32: return run_half<float, float, float>(p, half, c, s);
    case TW_BF16:
      if (p.coef16) return run_half<bf16, bf16, float>(p, half, c, s);
      return run_half<bf16, float, float>(p, half, c, s);
    default:
      if (p.coef16) return run_half<_Float16, _Float16, float>(p, half, c, s);
      return run_half<_Float16, float, float>(p, half, c, s);
  }
}

static hipError_t misc(const Plan& p, int what, const Pending* pend, hipStream_t s) {
  switch (p.precision) {
    case TW_FP64: return run_misc<double, double, double>(p, what, pend, s);
    case TW_FP32: return run_misc<float, float, float>(p, what, pend, s);
    case TW_BF16:
      if (p.coef16) return run_misc<bf16, bf16, float>(p, what, pend, s);
      return run_misc<bf16, float, float>(p, what, pend, s);
    default:
      if (p.coef16) return run_misc<_Float16, _Float16, float>(p, what, pend, s);
      return run_misc<_Float16, float, float>(p, what, pend, s);
  }
}

hipError_t launch_stress_energy(const Plan& p, hipStream_t s) { return misc(p, 0, nullptr, s); }
hipError_t launch_uhat(const Plan& p, hipStream_t s) { return misc(p, 1, nullptr, s); }
hipError_t launch_finalize(const Plan& p, const Pending& pend, hipStream_t s) { return misc(p, 2, &pend, s); }

hipError_t launch_check_finite(const Plan& p, int32_t* flag, hipStream_t s) {
  dim3 b(64, 4);
  for (int f = 0; f < NF; f++) {
    const DevField& d = p.f[f];
    if (!d.s) continue;
    dim3 g((d.cols + 63) / 64, (d.rows + 3) / 4);
    switch (p.precision) {
      case TW_FP64: {
        F<double> X{(double*)d.s, nullptr, nullptr, d.rows, d.cols, d.pitch};
        hipLaunchKernelGGL(k_finite<double>, g, b, 0, s, X, flag);
      } break;
      case TW_FP32: {
        F<float> X{(float*)d.s, nullptr, nullptr, d.rows, d.cols, d.pitch};
        hipLaunchKernelGGL(k_finite<float>, g, b, 0, s, X, flag);
      } break;
      case TW_BF16: {
        F<bf16> X{(bf16*)d.s, nullptr, nullptr, d.rows, d.cols, d.pitch};
        hipLaunchKernelGGL(k_finite<bf16>, g, b, 0, s, X, flag);
      } break;
      default: {
        F<_Float16> X{(_Float16*)d.s, nullptr, nullptr, d.rows, d.cols, d.pitch};
        hipLaunchKernelGGL(k_finite<_Float16>, g, b, 0, s, X, flag);
      } break;
    }
  }
  return hipGetLastError();
}

}  // namespace tw
