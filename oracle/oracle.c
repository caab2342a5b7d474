/*
 * TEST INFRASTRUCTURE ONLY: the CPU oracle for the fluid step (a C
 * restatement of the reference's algorithm; see oracle_impl.inc for the
 * per-function citations). Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it, always as the checker or the timed
 * CPU baseline, never as part of the product path.
 *
 * Build: make -C oracle   (gcc -O2 -ffp-contract=off, see Makefile)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static inline int64_t mini64(int64_t a, int64_t b) { return a < b ? a : b; }

/* Loop annotations of the all-cores build (liboracle_omp.so, -fopenmp): the
 * grid loops split their (z, y) rows over threads and the vector loops their
 * elements. Every element is still computed by the same expression, so only
 * the reductions (dot products, norms) may round differently; the serial
 * build (liboracle.so) ignores them. This build times the CPU path on all
 * host cores, standing in for the reference's rayon par_azip! loops
 * (panopaea/src/dec/grid.rs:277-333). */
#ifdef _OPENMP
#define OR_PAR_ROWS _Pragma("omp parallel for collapse(2) schedule(static)")
#define OR_PAR_VEC _Pragma("omp parallel for schedule(static)")
#define OR_PAR_SUM_S _Pragma("omp parallel for reduction(+:s) schedule(static)")
#define OR_PAR_SUM_RS2 _Pragma("omp parallel for reduction(+:rs2) schedule(static)")
#define OR_PAR_SUM_PQ _Pragma("omp parallel for reduction(+:pq) schedule(static)")
#define OR_PAR_SUM_M _Pragma("omp parallel for reduction(+:m) schedule(static)")
#else
#define OR_PAR_ROWS
#define OR_PAR_VEC
#define OR_PAR_SUM_S
#define OR_PAR_SUM_RS2
#define OR_PAR_SUM_PQ
#define OR_PAR_SUM_M
#endif

/* imgproc::transfer, panopaea_utils/src/imgproc.rs:2-5: x.min(upper).max(lower)
 * (f64::min/max return the other operand for a NaN, as fmin/fmax do), then
 * ((v - lower) / (upper - lower) * 255.0) as u8 (truncation). */
static inline unsigned char or_transfer(double x, double lower, double upper)
{
    double v = fmax(fmin(x, upper), lower);
    double t = (v - lower) / (upper - lower) * 255.0;
    return (unsigned char)(int)t;
}

/* Solve / step parameters shared by both scalar instantiations. */
typedef struct {
    int64_t d, h, w;        /* cells; d == 1 is 2D */
    int32_t bc[6];          /* x-, x+, y-, y+, z-, z+ : 0 wall, 1 open */
    double dt;
    int32_t precond;        /* 0 identity, 1 jacobi, 2 multigrid */
    int32_t levels, nu1, nu2;
    int32_t tol_kind;       /* 0 absolute L-inf (pcg.rs:58-60), 1 relative L-inf */
    double tol;
    int32_t max_iterations;
    int32_t mode;           /* 0 literal reference arithmetic, 1 HIP-path arithmetic */
    int32_t remove_mean;    /* subtract the mean of b first (closed boxes) */
    double buoyancy;        /* smoke buoyancy coefficient, 0 = none */
    double src_center[3];   /* (z, y, x) of the density source sphere */
    double src_radius;      /* <= 0: no source */
} or_params;

typedef struct {
    int32_t iterations;
    int32_t converged;
    double b_linf;
    double r_linf;          /* recursive residual at exit */
    double threshold;       /* absolute threshold used */
} or_info;

/* Threads the loops above run on: omp_get_max_threads() in the all-cores
 * build, 1 in the serial one. */
#ifdef _OPENMP
#include <omp.h>
int or_threads(void) { return omp_get_max_threads(); }
#else
int or_threads(void) { return 1; }
#endif

#define T float
#define FN(n) n##_f32
#include "oracle_impl.inc"
#include "oracle_drive.inc"
#undef T
#undef FN

#define T double
#define FN(n) n##_f64
#include "oracle_impl.inc"
#include "oracle_drive.inc"
#undef T
#undef FN
