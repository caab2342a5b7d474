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
