/*
 * tw_oracle.h -- CPU oracle for the staggered SBP-SAT wave-solver hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Results are checked against it: tests/, the smoke check and
 * bench.py's cpu_baseline leg load it; the product (libtwave.so) never links,
 * loads or calls it.
 *
 * It restates, loop by loop and in the reference's operation order, the CPU
 * algorithm of the reference:
 *   elastic:  cuda/src/cpp/grid.hpp, grid_drvt.tpp, grid_update.tpp,
 *             grid_energy.cpp, grid_parameter.cpp, forward_simulation.cpp,
 *             forward.cpp (the host twin of the CUDA solver);
 *   acoustic: singlecore/src/cpp/namespaces.hpp (ns_action, prepare_energy_weight),
 *             singlecore/src/main_cpp/cpu_acoustic.cpp (time loop).
 * and extends it, with the same per-point structure, to what the reference
 * lacks: SBP closures on the fast axis, characteristic (absorbing) SAT, and
 * 16-bit field storage with FP32 arithmetic.
 *
 * Pinning (tests/test_oracle_*.py):
 *   - acoustic periodic path: bitwise against the compiled reference
 *     (oracle/_ref, golden fixtures in tests/golden/);
 *   - SBP operators: exact SBP identity and polynomial exactness;
 *   - SAT free surface: discrete energy constant to round-off;
 *   - characteristic SAT: energy non-increasing, outgoing waves leave;
 *   - compensation variants: error-free-transformation identities;
 *   - bilinear interpolation, location maps, Ricker: closed forms / KATs.
 *
 * The configuration struct is tw_config from include/tw.h.
 */
#ifndef TW_ORACLE_H
#define TW_ORACLE_H
#include "../include/tw.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_solver or_solver;

or_solver* or_create(const tw_config* cfg);
void       or_destroy(or_solver* o);
const char* or_error(void);
int  or_set_model(or_solver* o, const double* rho, const double* vp, const double* vs,
                  int nxm, int nym, long num_dxm, long den_dxm);
int  or_set_coef(or_solver* o, int coef, const double* v);
int  or_get_coef(or_solver* o, int coef, double* v);
int  or_set_source(or_solver* o, int grid, int ix, int iy, const double* inc, int n);
int  or_set_source_ricker(or_solver* o, int grid, int ix, int iy, double f0, double t0);
int  or_set_receivers(or_solver* o, int n, const int* field, const int* ix, const int* iy);
int  or_step(or_solver* o, int nsteps);
int  or_read_receivers(or_solver* o, double* out, int* nt);
int  or_read_energy(or_solver* o, double* ek, double* ep0, double* ep1, int* nt);
int  or_get_field(or_solver* o, int field, double* out);
int  or_set_field(or_solver* o, int field, const double* in);
int  or_get_companion(or_solver* o, int field, double* out);
int  or_field_shape(const tw_config* cfg, int field, int* rows, int* cols);
int  or_locate(const tw_config* cfg, int grid, long num_x, long den_x, long num_y, long den_y,
               int* ix, int* iy);
int  or_interpolate(const tw_config* cfg, int grid, const double* model, int nxm, int nym,
                    long num_dxm, long den_dxm, double* out);
double or_ricker_increment(const tw_config* cfg, double f0, double t0, int it, int half);
/* operator tables after rounding to the arithmetic type, for the SBP tests:
 * which = 0 stencil[4], 1 D for N output left [4x5], 2 right [4x5],
 *         3 D for M output left [3x5], 4 right [3x5], 5 SAT lift N [L,R],
 *         6 SAT lift M L[3], 7 SAT lift M R[3], 8 H diag N [4], 9 H diag M [3] */
int  or_table(or_solver* o, int which, double* out);
void or_set_threads(int n);
/* the per-point update the time step applies, on n independent points:
 * FP64/FP32 reference variants (TW_CPST_0 .. TW_CPST_6C) update (s, r, c) in
 * place, r being the accumulated right-hand side (values must be exact in the
 * precision); BF16/FP16 store the FP32 sum t and return the stored value and
 * the K residual (value) or the K8 code; k8_decode maps codes to residuals */
int  or_update_points(int precision, int cpst, long n, double* s, double* r, double* c);
int  or_store16_points(int precision, int cpst, long n, const double* t, double* hi, double* res);
int  or_k8_decode_points(int precision, long n, const double* hi, const double* q, double* out);

#ifdef __cplusplus
}
#endif
#endif
