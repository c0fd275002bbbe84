/*
 * tw_oracle_body.h -- one precision instance of the CPU oracle (test
 * infrastructure only; see tw_oracle.h).  Included by tw_oracle.c with
 *   PFX     symbol prefix
 *   SREAL   field storage type       CREAL  arithmetic type
 *   S2C(x)  storage -> arithmetic    C2S(x) arithmetic -> storage (RNE)
 *   REFMODE 1 when CREAL == SREAL (reference compensation variants apply)
 */
#define OR_CAT2(a, b) a##_##b
#define OR_CAT(a, b) OR_CAT2(a, b)
#define FN(name) OR_CAT(PFX, name)

/* one multiply-accumulate a + x*c of the derivative sums and of their
 * products with the material coefficient (dx_term, dy_term, mm_term).
 * REFMODE (FP64, FP32): a separate product and sum, as the reference's host
 * code computes it.  16-bit storage: one fused multiply-add (fmaf, correctly
 * rounded), as nvcc contracts the reference's device loops
 * (cuda/src/cuda/grid_kernel_x.cut:38-40, v_d += S * c). */
#undef OR_MAC
#if REFMODE
#define OR_MAC(a, x, c) ((a) + (x) * (c))
#else
#define OR_MAC(a, x, c) fmaf((x), (c), (a))
#endif

typedef struct {
  int rows, cols;
  SREAL* s;   /* state                                   */
  SREAL* r;   /* R: persisting residual ('R' variants, K) */
  SREAL* c;   /* C: separate compensation ('C' variants)  */
  CREAL* acc; /* right-hand side accumulated this half step */
} FN(fld);

/* per-axis tap tables: for output index i on an output of type t along the axis */
typedef struct {
  int n;           /* output length along the axis          */
  int* ntap;       /* 4 or 5                                  */
  int* idx;        /* [n][5] partner indices                  */
  CREAL* cf;       /* [n][5] coefficients (already * dt/dx)   */
} FN(taps);

typedef struct {
  or_common* cm;
  FN(fld) f[6];               /* indexed by tw_field                          */
  CREAL* coef[6];             /* coefficient arrays in arithmetic type        */
  FN(taps) tx[2], ty[2];      /* [t] t = 0 for output type M, 1 for N         */
  CREAL st[4];                /* interior stencil * dt/dx                     */
  CREAL DN_L[4][5], DN_R[4][5];  /* output N, partner M: W_intr.DM_* * dt/dx   */
  CREAL DM_L[3][5], DM_R[3][5];  /* output M, partner N: W_intr.DN_* * dt/dx   */
  CREAL ainvN_L, ainvN_R;        /* SAT lift for N-type output: dt*1/(dx*7/18) */
  CREAL ainvM_L[3], ainvM_R[3];  /* SAT lift for M-type output                 */
  CREAL projM_L[3], projM_R[3];  /* extrapolation of M data to the boundary    */
  CREAL hN[4], hM[3];            /* H diagonal at the closure (A_diag * dx)     */
  CREAL* z[6][4];             /* characteristic impedance along each side, per velocity field */
} FN(st_t);

static void FN(free_taps)(FN(taps) * t) { free(t->ntap); free(t->idx); free(t->cf); }

/* Tap table along one axis.  out_t: 1 = output N-type, 0 = output M-type.
 * periodic: shifts {-2,-1,0,1} (N output) / {-1,0,1,2} (M output) taken mod M
 *   (cuda/src/cpp/grid.hpp:361-362, grid_drvt.tpp:45-167: N_modulo = M for both).
 * SBP: closure rows use W_intr (4x5 for N output, 3x5 for M output) on the first /
 *   last five partner points (grid_drvt.tpp:368-464), interior uses the shifts. */
static int FN(build_taps)(FN(st_t) * S, FN(taps) * T, int out_t, int M, int periodic) {
  int n = out_t ? M + 1 : M, np = out_t ? M : M + 1;
  T->n = n;
  T->ntap = (int*)calloc((size_t)n, sizeof(int));
  T->idx = (int*)calloc((size_t)n * 5, sizeof(int));
  T->cf = (CREAL*)calloc((size_t)n * 5, sizeof(CREAL));
  if (!T->ntap || !T->idx || !T->cf) return -1;
  const int shN[4] = {-2, -1, 0, 1}, shM[4] = {-1, 0, 1, 2};
  const int* sh = out_t ? shN : shM;
  int nb = out_t ? 4 : 3;
  for (int i = 0; i < n; i++) {
    int* ix = T->idx + 5 * i;
    CREAL* cf = T->cf + 5 * i;
    if (periodic) {
      T->ntap[i] = 4;
      for (int k = 0; k < 4; k++) { ix[k] = (i + sh[k] + M) % M; cf[k] = S->st[k]; }
    } else if (i < nb) {
      T->ntap[i] = 5;
      for (int w = 0; w < 5; w++) { ix[w] = w; cf[w] = out_t ? S->DN_L[i][w] : S->DM_L[i][w]; }
    } else if (i >= n - nb) {
      int r = i - (n - nb);
      T->ntap[i] = 5;
      for (int w = 0; w < 5; w++) { ix[w] = np - 5 + w; cf[w] = out_t ? S->DN_R[r][w] : S->DM_R[r][w]; }
    } else {
      T->ntap[i] = 4;
      for (int k = 0; k < 4; k++) { ix[k] = i + sh[k]; cf[k] = S->st[k]; }
    }
  }
  return 0;
}

static void FN(destroy)(void* p) {
  FN(st_t)* S = (FN(st_t)*)p;
  if (!S) return;
  for (int f = 0; f < 6; f++) {
    free(S->f[f].s); free(S->f[f].r); free(S->f[f].c); free(S->f[f].acc);
    free(S->coef[f]);
    for (int k = 0; k < 4; k++) free(S->z[f][k]);
  }
  for (int t = 0; t < 2; t++) { FN(free_taps)(&S->tx[t]); FN(free_taps)(&S->ty[t]); }
  free(S);
}

/* Operator tables: the reference builds them in FP64 and casts to the solver
 * precision (grid.hpp:274-464); SAT lifts are rounded twice (divide, cast,
 * then *= dt with dt first cast, container_run_time.hpp:124). */
static void FN(tables)(FN(st_t) * S) {
  const tw_config* c = &S->cm->cfg;
  double dt = c->dt, dx = c->dx;
  for (int k = 0; k < 4; k++)
    S->st[k] = c->stencil_form ? (CREAL)(OR_STENCIL[k] * dt / dx) : (CREAL)(OR_STENCIL[k] * (dt / dx));
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 5; j++) {
      S->DN_L[i][j] = (CREAL)(OR_DM_L[i][j] * (dt / dx));
      S->DN_R[i][j] = (CREAL)(OR_DM_R[i][j] * (dt / dx));
    }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 5; j++) {
      S->DM_L[i][j] = (CREAL)(OR_DN_L[i][j] * (dt / dx));
      S->DM_R[i][j] = (CREAL)(OR_DN_R[i][j] * (dt / dx));
    }
  CREAL dtc = (CREAL)dt;
  S->ainvN_L = (CREAL)((double)(CREAL)(1.0 / (dx * OR_AN_L[0])) * (double)dtc);
  S->ainvN_R = (CREAL)((double)(CREAL)(1.0 / (dx * OR_AN_R[3])) * (double)dtc);
  for (int i = 0; i < 3; i++) {
    S->ainvM_L[i] = (CREAL)((double)(CREAL)(OR_PM_L[i] / (dx * OR_AM_L[i])) * (double)dtc);
    S->ainvM_R[i] = (CREAL)((double)(CREAL)(OR_PM_R[i] / (dx * OR_AM_R[i])) * (double)dtc);
    S->projM_L[i] = (CREAL)OR_PM_L[i];
    S->projM_R[i] = (CREAL)OR_PM_R[i];
  }
  for (int i = 0; i < 4; i++) S->hN[i] = (CREAL)(OR_AN_L[i] * dx);
  for (int i = 0; i < 3; i++) S->hM[i] = (CREAL)(OR_AM_L[i] * dx);
}

static int FN(alloc_field)(FN(st_t) * S, int f) {
  const tw_config* c = &S->cm->cfg;
  FN(fld)* F = &S->f[f];
  F->rows = S->cm->rows[f];
  F->cols = S->cm->cols[f];
  size_t n = (size_t)F->rows * F->cols;
  F->s = (SREAL*)calloc(n, sizeof(SREAL));
  F->r = (SREAL*)calloc(n, sizeof(SREAL));
  F->acc = (CREAL*)calloc(n, sizeof(CREAL));
  if (c->cpst == TW_CPST_3C || c->cpst == TW_CPST_M3C || c->cpst == TW_CPST_6C)
    F->c = (SREAL*)calloc(n, sizeof(SREAL));
  return (F->s && F->r && F->acc) ? 0 : -1;
}

static void* FN(create)(or_common* cm) {
  FN(st_t)* S = (FN(st_t)*)calloc(1, sizeof(FN(st_t)));
  if (!S) return NULL;
  S->cm = cm;
  FN(tables)(S);
  const tw_config* c = &cm->cfg;
  int perx = c->bc[0] == 'P', pery = c->bc[2] == 'P';
  for (int t = 0; t < 2; t++) {
    if (FN(build_taps)(S, &S->tx[t], t, c->mx, perx) || FN(build_taps)(S, &S->ty[t], t, c->my, pery)) {
      FN(destroy)(S);
      return NULL;
    }
  }
  for (int k = 0; k < cm->nf; k++)
    if (FN(alloc_field)(S, cm->fields[k])) { FN(destroy)(S); return NULL; }
  return S;
}

/* coefficient arrays arrive as doubles already rounded to their storage type */
static void FN(load_coef)(void* p, int k) {
  FN(st_t)* S = (FN(st_t)*)p;
  or_common* cm = S->cm;
  size_t n = (size_t)cm->coef_rows[k] * cm->coef_cols[k];
  if (!S->coef[k]) S->coef[k] = (CREAL*)malloc(n * sizeof(CREAL));
  for (size_t i = 0; i < n; i++) S->coef[k][i] = (CREAL)cm->coef[k][i];
}

static void FN(load_z)(void* p) {
  FN(st_t)* S = (FN(st_t)*)p;
  or_common* cm = S->cm;
  for (int f = 0; f < 6; f++)
    for (int side = 0; side < 4; side++) {
      if (!cm->z[f][side]) continue;
      int n = cm->zlen[f][side];
      if (!S->z[f][side]) S->z[f][side] = (CREAL*)malloc((size_t)n * sizeof(CREAL));
      for (int i = 0; i < n; i++) S->z[f][side][i] = (CREAL)cm->z[f][side][i];
    }
}

static void FN(set_field)(void* p, int f, const double* in, int which) {
  FN(st_t)* S = (FN(st_t)*)p;
  FN(fld)* F = &S->f[f];
  size_t n = (size_t)F->rows * F->cols;
  SREAL* d = which == 0 ? F->s : (which == 1 ? F->r : F->c);
  if (!d) return;
  for (size_t i = 0; i < n; i++) d[i] = C2S((CREAL)in[i]);
}

static void FN(get_field)(void* p, int f, double* out, int which) {
  FN(st_t)* S = (FN(st_t)*)p;
  FN(fld)* F = &S->f[f];
  size_t n = (size_t)F->rows * F->cols;
  SREAL* d = which == 0 ? F->s : (which == 1 ? F->r : F->c);
  for (size_t i = 0; i < n; i++) out[i] = d ? (double)S2C(d[i]) : 0.0;
}

#if !REFMODE
/* K8: the residual array holds an integer code q in [-127, 127] (exact in the
 * storage type); the residual is q * ulp(hi) / 256 of the stored value hi */
static inline CREAL FN(k8_dec)(SREAL q, SREAL hi) {
  const int e = S_EXP(hi);
  return ldexpf(S2C(q), (e > 0 ? e : 1) - K8_BIAS);
}
static inline SREAL FN(k8_enc)(CREAL x, SREAL hi) {
  const int e = S_EXP(hi);
  float q = rintf(ldexpf(x, K8_BIAS - (e > 0 ? e : 1)));
  q = q > 127.f ? 127.f : (q < -127.f ? -127.f : q);
  return C2S(q);
}
#endif

/* value of a stored point as seen by receivers and energy: the state, plus the
 * Kahan residual in the 16-bit K (kmode 1) and K8 (kmode 2) variants */
static inline CREAL FN(val)(const FN(fld) * F, size_t i, int kmode) {
  CREAL v = S2C(F->s[i]);
#if !REFMODE
  if (kmode == 2) return v + FN(k8_dec)(F->r[i], F->s[i]);
#endif
  if (kmode) v = v + S2C(F->r[i]);
  return v;
}

/* a value as the stencils, SAT terms and boundary values read it: the stored
 * value, or for K8S the stored value plus its decoded residual (FP32 sum) */
static inline CREAL FN(pv)(const FN(fld) * F, size_t i, int rf) {
#if !REFMODE
  if (rf) return S2C(F->s[i]) + FN(k8_dec)(F->r[i], F->s[i]);
#endif
  (void)rf;
  return S2C(F->s[i]);
}

/* ---------- right-hand side accumulation ---------- */

/* start of a half step: acc = R (reference variants; R is zero after the
 * '0' and 'C' updates) or the stored Kahan residual / zero (16-bit modes) */
static void FN(acc_init)(FN(st_t) * S, int f) {
  FN(fld)* F = &S->f[f];
  size_t n = (size_t)F->rows * F->cols;
#if REFMODE
  for (size_t i = 0; i < n; i++) F->acc[i] = F->r[i];
#else
  int k = S->cm->cfg.cpst == TW_CPST_K, k8 = OR_K8(S->cm->cfg.cpst);
  for (size_t i = 0; i < n; i++)
    F->acc[i] = k ? S2C(F->r[i]) : (k8 ? FN(k8_dec)(F->r[i], F->s[i]) : (CREAL)0);
#endif
}

/* x-derivative of partner field pf into acc of field f, times coefficient
 * (grid_drvt.tpp:368-464 for SBP rows, :45-167 structure for periodic; the
 * per-point order is v = 0; v = v + S*c per tap; R = R + v*P, drvt_to_rths
 * grid_drvt.tpp:23-34; each step one OR_MAC).  extra: MM elastic, the same v also feeds the second
 * stress with lambda and the first (dir 0) / second (dir 1) with 2 mu. */
static void FN(dx_term)(FN(st_t) * S, int f, int pf, int coefk, int out_t) {
  FN(fld)* F = &S->f[f];
  FN(fld)* P = &S->f[pf];
  const CREAL* cp = S->coef[coefk];
  FN(taps)* T = &S->tx[out_t];
  int rows = F->rows, cols = F->cols;
  const int rf = S->cm->cfg.cpst == TW_CPST_K8S;
#pragma omp parallel for schedule(static)
  for (int ix = 0; ix < rows; ix++) {
    int nt = T->ntap[ix];
    const int* id = T->idx + 5 * ix;
    const CREAL* cf = T->cf + 5 * ix;
    CREAL* a = F->acc + (size_t)ix * cols;
    const CREAL* pr = cp + (size_t)ix * cols;
    for (int iy = 0; iy < cols; iy++) {
      CREAL v = 0;
      for (int k = 0; k < nt; k++) v = OR_MAC(v, FN(pv)(P, (size_t)id[k] * P->cols + iy, rf), cf[k]);
      a[iy] = OR_MAC(a[iy], v, pr[iy]);
    }
  }
}

static void FN(dy_term)(FN(st_t) * S, int f, int pf, int coefk, int out_t) {
  FN(fld)* F = &S->f[f];
  FN(fld)* P = &S->f[pf];
  const CREAL* cp = S->coef[coefk];
  FN(taps)* T = &S->ty[out_t];
  int rows = F->rows, cols = F->cols;
  const int rf = S->cm->cfg.cpst == TW_CPST_K8S;
#pragma omp parallel for schedule(static)
  for (int ix = 0; ix < rows; ix++) {
    CREAL* a = F->acc + (size_t)ix * cols;
    const CREAL* pr = cp + (size_t)ix * cols;
    const size_t p0 = (size_t)ix * P->cols;
    for (int iy = 0; iy < cols; iy++) {
      int nt = T->ntap[iy];
      const int* id = T->idx + 5 * iy;
      const CREAL* cf = T->cf + 5 * iy;
      CREAL v = 0;
      for (int k = 0; k < nt; k++) v = OR_MAC(v, FN(pv)(P, p0 + id[k], rf), cf[k]);
      a[iy] = OR_MAC(a[iy], v, pr[iy]);
    }
  }
}

/* MM elastic: sxx/syy share each derivative (drvt_to_rths<true>) */
static void FN(mm_term)(FN(st_t) * S, int pf, int dir) {
  FN(fld)* A = &S->f[TW_SXX];
  FN(fld)* B = &S->f[TW_SYY];
  FN(fld)* P = &S->f[pf];
  const CREAL* lam = S->coef[TW_C_LAMBDA];
  const CREAL* tmu = S->coef[TW_C_TWOMU];
  int rows = A->rows, cols = A->cols;
  FN(taps)* T = dir == 0 ? &S->tx[0] : &S->ty[0];
  const int rf = S->cm->cfg.cpst == TW_CPST_K8S;
#pragma omp parallel for schedule(static)
  for (int ix = 0; ix < rows; ix++) {
    for (int iy = 0; iy < cols; iy++) {
      int i = dir == 0 ? ix : iy;
      int nt = T->ntap[i];
      const int* id = T->idx + 5 * i;
      const CREAL* cf = T->cf + 5 * i;
      CREAL v = 0;
      if (dir == 0)
        for (int k = 0; k < nt; k++) v = OR_MAC(v, FN(pv)(P, (size_t)id[k] * P->cols + iy, rf), cf[k]);
      else
        for (int k = 0; k < nt; k++) v = OR_MAC(v, FN(pv)(P, (size_t)ix * P->cols + id[k], rf), cf[k]);
      size_t q = (size_t)ix * cols + iy;
      A->acc[q] = OR_MAC(A->acc[q], v, lam[q]);
      B->acc[q] = OR_MAC(B->acc[q], v, lam[q]);
      if (dir == 0) A->acc[q] = OR_MAC(A->acc[q], v, tmu[q]);
      else          B->acc[q] = OR_MAC(B->acc[q], v, tmu[q]);
    }
  }
}

/* SAT on a velocity field along one axis (secure_bdry, grid_drvt.tpp:469-695):
 *   SEND = sum_w proj(w) * S_partner(w)   (accumulated from 0)
 *   RECV = 0 (free surface) | Z*u_hat (left) | -(Z*u_hat) (right)   [characteristic]
 *   v_d  = eta * Ainv(w) * (SEND - RECV);  R(w) = R(w) + v_d * P(w)
 * eta = +1 left, -1 right.  u_hat is the field's own boundary value (N type
 * along the axis) or its extrapolation with projM (M type).  For the
 * characteristic case SEND - RECV = traction -/+ Z*u imposes the incoming
 * characteristic to zero; the free surface is the reference's case. */
static void FN(sat)(FN(st_t) * S, int f, int pf, int coefk, int axis, int out_t) {
  const tw_config* c = &S->cm->cfg;
  FN(fld)* F = &S->f[f];
  FN(fld)* P = &S->f[pf];
  const CREAL* cp = S->coef[coefk];
  const int rf = c->cpst == TW_CPST_K8S;
  for (int side = 0; side < 2; side++) {
    char bc = c->bc[2 * axis + side];
    if (bc == 'P') continue;
    int chr = bc == 'C';
    int n_along = axis == 0 ? F->rows : F->cols;      /* length along the axis (own)     */
    int np_along = axis == 0 ? P->rows : P->cols;     /* partner length along the axis   */
    int n_across = axis == 0 ? F->cols : F->rows;
    int plen = out_t ? 3 : 1;                         /* partner is M type if out is N   */
    int alen = out_t ? 1 : 3;
    CREAL eta = side == 0 ? (CREAL)1 : (CREAL)-1;
    const CREAL* proj = out_t ? (side == 0 ? S->projM_L : S->projM_R) : NULL;
    const CREAL* z = chr ? S->z[f][2 * axis + side] : NULL;
    for (int j = 0; j < n_across; j++) {
      CREAL send = 0;
      for (int w = 0; w < plen; w++) {
        int pa = side == 0 ? w : np_along - plen + w;
        size_t q = axis == 0 ? (size_t)pa * P->cols + j : (size_t)j * P->cols + pa;
        CREAL pw = out_t ? proj[w] : (CREAL)1;
        send = send + pw * FN(pv)(P, q, rf);
      }
      CREAL recv = 0;
      if (chr) {
        CREAL u = 0;
        if (out_t) {
          int a = side == 0 ? 0 : n_along - 1;
          size_t q = axis == 0 ? (size_t)a * F->cols + j : (size_t)j * F->cols + a;
          u = FN(pv)(F, q, rf);
        } else {
          const CREAL* pm = side == 0 ? S->projM_L : S->projM_R;
          for (int w = 0; w < 3; w++) {
            int a = side == 0 ? w : n_along - 3 + w;
            size_t q = axis == 0 ? (size_t)a * F->cols + j : (size_t)j * F->cols + a;
            u = u + pm[w] * FN(pv)(F, q, rf);
          }
        }
        CREAL zu = z[j] * u;
        recv = side == 0 ? zu : -zu;
      }
      CREAL g = send - recv;
      for (int w = 0; w < alen; w++) {
        int a = side == 0 ? w : n_along - alen + w;
        CREAL ai = out_t ? (side == 0 ? S->ainvN_L : S->ainvN_R) : (side == 0 ? S->ainvM_L[w] : S->ainvM_R[w]);
        size_t q = axis == 0 ? (size_t)a * F->cols + j : (size_t)j * F->cols + a;
        CREAL vd = (eta * ai) * g;
        F->acc[q] = F->acc[q] + vd * cp[q];
      }
    }
  }
}

/* ---------- update (grid_update.tpp:1-128, grid_update.cut:28-124) ---------- */
#if REFMODE
/* one point of the leapfrog update S <- S + R in the reference's variants.
 * In: s (state), r (accumulated right-hand side, including the residual the
 * 'R' variants carried over), c (the 'C' variants' separate compensation).
 * '0': S = s + r, R = 0 (device_update_00, grid_update.cut:62-67);
 * 3R: Fast2Sum (s, r) -> (S, R): a = s + r, z = a - s, R = r - z
 *     (device_fast2sum :28-34 via device_update_3R_fast2sum :113-117;
 *      sum_3op, singlecore/src/cpp/namespaces.hpp:633-640);
 * -3R: the same three operations written out (device_update_3R :102-110);
 * 6R: 2Sum (s, r) -> (S, R): a = s + r, pa = a - r, pb = a - pa,
 *     R = (s - pa) + (r - pb) (device_slow2sum :42-55 via
 *     device_update_6R_slow2sum :120-124; sum_6op, namespaces.hpp:643-655);
 * 3C / -3C / 6C: the same with Y = r + c in place of r, the error into C
 *     and R = 0 (device_update_3C_fast2sum :83-89, device_update_3C :70-80,
 *     device_update_6C_slow2sum :92-98).  The dispatch on (cpst_N, cpst_S)
 *     is grid_update.cut:143-151. */
static inline void FN(upd_point)(int cpst, SREAL* ps, SREAL* pr, SREAL* pc, CREAL r) {
  CREAL s = *ps;
  switch (cpst) {
    case TW_CPST_0: *ps = s + r; *pr = 0; break;
    case TW_CPST_3R: { CREAL a = s + r; CREAL z = a - s; *pr = r - z; *ps = a; } break;
    case TW_CPST_M3R: { CREAL Y = r; CREAL T = s + Y; CREAL q = T - s; q = Y - q; *pr = q; *ps = T; } break;
    case TW_CPST_6R: {
      CREAL a = s + r; CREAL pa = a - r; CREAL pb = a - pa; CREAL da = s - pa; CREAL db = r - pb;
      *ps = a; *pr = da + db;
    } break;
    case TW_CPST_3C: { CREAL Y = r + *pc; CREAL a = s + Y; CREAL z = a - s; *pc = Y - z; *ps = a; *pr = 0; } break;
    case TW_CPST_M3C: { CREAL Y = r + *pc; CREAL T = s + Y; CREAL q = T - s; q = Y - q; *pc = q; *ps = T; *pr = 0; } break;
    case TW_CPST_6C: {
      CREAL Y = r + *pc; CREAL a = s + Y; CREAL pa = a - Y; CREAL pb = a - pa; CREAL da = s - pa; CREAL db = Y - pb;
      *ps = a; *pc = da + db; *pr = 0;
    } break;
  }
}
#else
/* one point of the 16-bit-storage update: FP32 sum t = s + acc, stored
 * S = RN16(t); K keeps RN16(t - S), K8 the code of t - S */
static inline void FN(upd_point16)(int cpst, SREAL* ps, SREAL* pr, CREAL acc) {
  CREAL t = S2C(*ps) + acc;
  SREAL hi = C2S(t);
  *ps = hi;
  if (cpst == TW_CPST_K) *pr = C2S(t - S2C(hi));
  if (OR_K8(cpst)) *pr = FN(k8_enc)(t - S2C(hi), hi);
}
#endif

static void FN(update)(FN(st_t) * S, int f) {
  FN(fld)* F = &S->f[f];
  size_t n = (size_t)F->rows * F->cols;
  int cpst = S->cm->cfg.cpst;
  SREAL zero = 0;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)n; i++) {
#if REFMODE
    SREAL cdummy = zero;
    FN(upd_point)(cpst, &F->s[i], &F->r[i], F->c ? &F->c[i] : &cdummy, F->acc[i]);
#else
    SREAL rdummy = zero;
    FN(upd_point16)(cpst, &F->s[i], F->r ? &F->r[i] : &rdummy, F->acc[i]);
#endif
  }
}

/* point source on every field of the source grid (forward_simulation.cpp:57-72,
 * 99-114): S = (T)((double)S + increment) */
static void FN(apply_source)(FN(st_t) * S, int half) {
  or_common* cm = S->cm;
  if (cm->src_grid < 0 || !cm->src_inc || cm->it >= cm->src_n) return;
  int g = cm->src_grid;
  int vgrid = (g == TW_GRID_MN || g == TW_GRID_NM);
  if ((half == 0) != vgrid) return;
  double inc = cm->src_inc[cm->it];
  for (int k = 0; k < cm->nf; k++) {
    int f = cm->fields[k];
    if (or_field_grid(cm->cfg.physics, f) != g) continue;
    FN(fld)* F = &S->f[f];
    size_t q = (size_t)cm->src_ix * F->cols + cm->src_iy;
#if REFMODE
    F->s[q] = (SREAL)((double)F->s[q] + inc);
#else
    if (cm->cfg.cpst == TW_CPST_K) {
      double v = (double)S2C(F->s[q]) + (double)S2C(F->r[q]) + inc;
      CREAL vf = (CREAL)v;
      SREAL hi = C2S(vf);
      F->s[q] = hi;
      F->r[q] = C2S(vf - S2C(hi));
    } else if (OR_K8(cm->cfg.cpst)) {
      double v = (double)S2C(F->s[q]) + (double)FN(k8_dec)(F->r[q], F->s[q]) + inc;
      CREAL vf = (CREAL)v;
      SREAL hi = C2S(vf);
      F->s[q] = hi;
      F->r[q] = FN(k8_enc)(vf - S2C(hi), hi);
    } else {
      F->s[q] = C2S((CREAL)((double)S2C(F->s[q]) + inc));
    }
#endif
  }
}

/* ---------- energy (grid_energy.cpp:180-216; cpu_acoustic.cpp:327-344) ---------- */
static double FN(energy_field)(FN(st_t) * S, int f) {
  or_common* cm = S->cm;
  const tw_config* c = &cm->cfg;
  FN(fld)* F = &S->f[f];
  int kmode = c->cpst == TW_CPST_K ? 1 : (OR_K8(c->cpst) ? 2 : 0);
  int g = or_field_grid(c->physics, f);
  int tx = (g == TW_GRID_NN || g == TW_GRID_NM), ty = (g == TW_GRID_NN || g == TW_GRID_MN);
  double E = 0.0;
  if (c->energy_form == 1) {
    /* acoustic reference: soln*prmt*soln*w, prmt = update coefficient, w = dx^2/2,
       halved at both ends of a node axis (namespaces.hpp:734-777) */
    const CREAL* prm = S->coef[or_field_coef(c->physics, f)];
    double w0 = c->dx * c->dx / 2.;
    for (int ix = 0; ix < F->rows; ix++)
      for (int iy = 0; iy < F->cols; iy++) {
        size_t q = (size_t)ix * F->cols + iy;
        double w = w0;
        if (ty && (iy == 0 || iy == F->cols - 1)) w /= 2.;
        if (tx && (ix == 0 || ix == F->rows - 1)) w /= 2.;
        CREAL s = FN(val)(F, q, kmode);
        E += (double)((CREAL)((CREAL)(s * prm[q]) * s)) * w;
      }
    return E;
  }
  /* H-norm: weight = compliance * hx(ix) * hy(iy) (adjust_parameters_energy_periodic) */
  for (int ix = 0; ix < F->rows; ix++) {
    double hx = or_hweight(c, 0, tx, ix, F->rows, S->hN, S->hM);
    for (int iy = 0; iy < F->cols; iy++) {
      double hy = or_hweight(c, 1, ty, iy, F->cols, S->hN, S->hM);
      size_t q = (size_t)ix * F->cols + iy;
      double s = (double)FN(val)(F, q, kmode);
      double comp;
      if (f == TW_VX || f == TW_VY) comp = 1. / (double)S->coef[or_field_coef(c->physics, f)][q];
      else if (f == TW_SXY) comp = 1. / (double)S->coef[TW_C_MU][q];
      else if (f == TW_P) comp = 1. / (double)S->coef[TW_C_KAPPA][q];
      else comp = 0.;
      double w = (comp * hx) * hy;
      E = E + s * w * s;
    }
  }
  return E / 2.;
}

static double FN(energy_mm)(FN(st_t) * S) {
  or_common* cm = S->cm;
  const tw_config* c = &cm->cfg;
  FN(fld)* A = &S->f[TW_SXX];
  FN(fld)* B = &S->f[TW_SYY];
  int kmode = c->cpst == TW_CPST_K ? 1 : (OR_K8(c->cpst) ? 2 : 0);
  const CREAL* lam = S->coef[TW_C_LAMBDA];
  const CREAL* tmu = S->coef[TW_C_TWOMU];
  double E = 0.0;
  for (int ix = 0; ix < A->rows; ix++) {
    double hx = or_hweight(c, 0, 0, ix, A->rows, S->hN, S->hM);
    for (int iy = 0; iy < A->cols; iy++) {
      double hy = or_hweight(c, 1, 0, iy, A->cols, S->hN, S->hM);
      size_t q = (size_t)ix * A->cols + iy;
      double L = (double)lam[q], M2 = (double)tmu[q];
      double p1 = (L + M2) / (M2 * (2. * L + M2));
      double p2 = -(L * 2.) / (M2 * (2. * L + M2));
      p1 = (p1 * hx) * hy;
      p2 = (p2 * hx) * hy;
      double sxx = (double)FN(val)(A, q, kmode), syy = (double)FN(val)(B, q, kmode);
      E = E + sxx * p1 * sxx;
      E = E + syy * p1 * syy;
      E = E + sxx * p2 * syy;
    }
  }
  return E / 2.;
}

static double FN(energy_stress)(FN(st_t) * S) {
  if (S->cm->cfg.physics == TW_ELASTIC) return FN(energy_mm)(S) + FN(energy_field)(S, TW_SXY);
  return FN(energy_field)(S, TW_P);
}

static double FN(energy_velocity)(FN(st_t) * S) {
  if (S->cm->cfg.physics == TW_ELASTIC) return FN(energy_field)(S, TW_VY) + FN(energy_field)(S, TW_VX);
  return FN(energy_field)(S, TW_VX) + FN(energy_field)(S, TW_VY);
}

/* ---------- one leapfrog step (forward_simulation.cpp:30-179) ---------- */
static void FN(one_step)(void* p) {
  FN(st_t)* S = (FN(st_t)*)p;
  or_common* cm = S->cm;
  const tw_config* c = &cm->cfg;
  int it = cm->it;
  int rec = c->record_energy && it < c->nt_max;
  if (rec) {
    if (!cm->ep_valid) { cm->ep_prev = FN(energy_stress)(S); cm->ep_valid = 1; }
    cm->ep0[it] = cm->ep_prev;
  }
  if (c->physics == TW_ELASTIC) {
    /* velocity half step: vy ("Vx", MN) then vx ("Vy", NM) */
    FN(acc_init)(S, TW_VY);
    FN(dx_term)(S, TW_VY, TW_SXY, TW_C_BINV_VY, 0);
    FN(sat)(S, TW_VY, TW_SXY, TW_C_BINV_VY, 0, 0);
    FN(dy_term)(S, TW_VY, TW_SYY, TW_C_BINV_VY, 1);
    FN(sat)(S, TW_VY, TW_SYY, TW_C_BINV_VY, 1, 1);
    FN(acc_init)(S, TW_VX);
    FN(dx_term)(S, TW_VX, TW_SXX, TW_C_BINV_VX, 1);
    FN(sat)(S, TW_VX, TW_SXX, TW_C_BINV_VX, 0, 1);
    FN(dy_term)(S, TW_VX, TW_SXY, TW_C_BINV_VX, 0);
    FN(sat)(S, TW_VX, TW_SXY, TW_C_BINV_VX, 1, 0);
    FN(update)(S, TW_VY);
    FN(update)(S, TW_VX);
    FN(apply_source)(S, 0);
    if (rec) cm->ek[it] = FN(energy_velocity)(S);
    /* stress half step: sxy (NN), then sxx/syy (MM) */
    FN(acc_init)(S, TW_SXY);
    FN(dx_term)(S, TW_SXY, TW_VY, TW_C_MU, 1);
    FN(dy_term)(S, TW_SXY, TW_VX, TW_C_MU, 1);
    FN(acc_init)(S, TW_SXX);
    FN(acc_init)(S, TW_SYY);
    FN(mm_term)(S, TW_VX, 0);
    FN(mm_term)(S, TW_VY, 1);
    FN(update)(S, TW_SXY);
    FN(update)(S, TW_SXX);
    FN(update)(S, TW_SYY);
  } else {
    /* acoustic (cpu_acoustic.cpp:231-322): vx from dp/dx, vy from dp/dy, p */
    FN(acc_init)(S, TW_VX);
    FN(dx_term)(S, TW_VX, TW_P, TW_C_BINV_VX, 1);
    FN(sat)(S, TW_VX, TW_P, TW_C_BINV_VX, 0, 1);
    FN(update)(S, TW_VX);
    FN(acc_init)(S, TW_VY);
    FN(dy_term)(S, TW_VY, TW_P, TW_C_BINV_VY, 1);
    FN(sat)(S, TW_VY, TW_P, TW_C_BINV_VY, 1, 1);
    FN(update)(S, TW_VY);
    FN(apply_source)(S, 0);
    if (rec) cm->ek[it] = FN(energy_velocity)(S);
    FN(acc_init)(S, TW_P);
    FN(dx_term)(S, TW_P, TW_VX, TW_C_KAPPA, 0);
    FN(dy_term)(S, TW_P, TW_VY, TW_C_KAPPA, 0);
    FN(update)(S, TW_P);
  }
  FN(apply_source)(S, 1);
  if (rec) {
    cm->ep_prev = FN(energy_stress)(S);
    cm->ep1[it] = cm->ep_prev;
  }
  if (it < c->nt_max) {
    int kmode = c->cpst == TW_CPST_K ? 1 : (OR_K8(c->cpst) ? 2 : 0);
    for (int r = 0; r < cm->nrcv; r++) {
      FN(fld)* F = &S->f[cm->rf[r]];
      size_t q = (size_t)cm->rix[r] * F->cols + cm->riy[r];
      cm->rec[(size_t)r * c->nt_max + it] = (double)FN(val)(F, q, kmode);
    }
  }
  cm->it++;
}

static int FN(table)(void* p, int which, double* out) {
  FN(st_t)* S = (FN(st_t)*)p;
  switch (which) {
    case 0: for (int k = 0; k < 4; k++) out[k] = S->st[k]; return 4;
    case 1: for (int k = 0; k < 20; k++) out[k] = S->DN_L[k / 5][k % 5]; return 20;
    case 2: for (int k = 0; k < 20; k++) out[k] = S->DN_R[k / 5][k % 5]; return 20;
    case 3: for (int k = 0; k < 15; k++) out[k] = S->DM_L[k / 5][k % 5]; return 15;
    case 4: for (int k = 0; k < 15; k++) out[k] = S->DM_R[k / 5][k % 5]; return 15;
    case 5: out[0] = S->ainvN_L; out[1] = S->ainvN_R; return 2;
    case 6: for (int k = 0; k < 3; k++) out[k] = S->ainvM_L[k]; return 3;
    case 7: for (int k = 0; k < 3; k++) out[k] = S->ainvM_R[k]; return 3;
    case 8: for (int k = 0; k < 4; k++) out[k] = S->hN[k]; return 4;
    case 9: for (int k = 0; k < 3; k++) out[k] = S->hM[k]; return 3;
  }
  return -1;
}

/* single-point entry points of the update, for the tests of the
 * error-free transformations and of the residual codes (the same functions
 * the time step calls) */
#if REFMODE
static void FN(update_points)(int cpst, long n, double* s, double* r, double* c) {
  for (long i = 0; i < n; i++) {
    SREAL ss = (SREAL)s[i], rr = (SREAL)r[i], cc = (SREAL)c[i];
    FN(upd_point)(cpst, &ss, &rr, &cc, (CREAL)r[i]);
    s[i] = (double)ss;
    r[i] = (double)rr;
    c[i] = (double)cc;
  }
}
#else
static void FN(store_points)(int cpst, long n, const double* t, double* hi, double* res) {
  for (long i = 0; i < n; i++) {
    SREAL ss = C2S((CREAL)0), rr = C2S((CREAL)0);
    FN(upd_point16)(cpst, &ss, &rr, (CREAL)t[i]);
    hi[i] = (double)S2C(ss);
    res[i] = (double)S2C(rr);
  }
}
static void FN(k8_decode_points)(long n, const double* hi, const double* q, double* out) {
  for (long i = 0; i < n; i++) out[i] = (double)FN(k8_dec)(C2S((CREAL)q[i]), C2S((CREAL)hi[i]));
}
#endif

static const or_vtbl FN(vt) = {FN(create), FN(destroy), FN(load_coef), FN(load_z), FN(set_field),
                               FN(get_field), FN(one_step), FN(table)};

#undef FN
#undef OR_CAT
#undef OR_CAT2
