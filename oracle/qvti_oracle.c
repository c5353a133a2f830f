/*
 * qvti_oracle.c -- CPU oracle translation unit (TEST INFRASTRUCTURE ONLY).
 * Instantiates qvti_oracle_impl.h in FP32 and FP64.  Loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
 * product path.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "qvti_oracle.h"

/* The reference's truncated pi (cuviscoelas_3D.cu:15, viscoelas_3D.cpp:14). */
#define QO_PI 3.1415926

/* Q-model coefficients of one cell, cuviscoelas_3D.cu:195-205 (tao_*1, tao_*2,
 * eta_*1..3), evaluated left to right in FP64 as the reference does; w0 is
 * the FP32 value the host passes (viscoelas_3D.cpp:109). */
void qo_q_coeffs(double rho, double v, double gamma, double w0, double *out)
{
    double c = v * cos(gamma * QO_PI / 2.0);
    out[0] = -rho * pow(c, 1) * gamma * w0;                    /* eta1 */
    out[1] = rho * pow(c, 2);                                  /* eta2 */
    out[2] = rho * pow(c, 3) * gamma / w0;                     /* eta3 */
    out[3] = rho * pow(c, 1) * gamma * QO_PI;                  /* tau1 */
    out[4] = rho * pow(c, 2) * pow(gamma, 2) * QO_PI / w0;     /* tau2 */
}

#define REAL float
#define SFX _f32
#define QO_IS_F32 1
#include "qvti_oracle_impl.h"
#undef REAL
#undef SFX
#undef QO_IS_F32

#define REAL double
#define SFX _f64
#define QO_IS_F32 0
#include "qvti_oracle_impl.h"
#undef REAL
#undef SFX
#undef QO_IS_F32

int qo_version(void) { return 1; }
