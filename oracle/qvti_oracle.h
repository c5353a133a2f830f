/*
 * qvti_oracle.h -- CPU oracle for the Q-VTI time-stepping loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a literal CPU restatement of the
 * reference's per-shot loop (cuviscoelas_3D.cu:1255-1605), against which the
 * HIP engine's results are checked.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product path never does.
 *
 * It keeps the reference's structure on purpose: nine 3-D C2C forward
 * transforms and 54 3-D C2C inverse transforms per step, `.x / ntp` on every
 * inverse output, the same update formulas and the same quirks (truncated pi,
 * velp_max = 5000 in the C-PML, last-plane skip, 8-point shear averages,
 * backward-difference dt-terms on the pre-injection spectrum).  The FFT is
 * this file's own mixed-radix Stockham transform (cuFFT is not available).
 *
 * Two builds of the same source: *_f32 (FP32 fields and FFT, tables and
 * coefficients rounded to FP32 as the reference stores them) and *_f64 (the
 * same scheme entirely in FP64: the error floor).
 */
#ifndef QVTI_ORACLE_H
#define QVTI_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Source kinds.  QO_SRC_FX is the reference default (cuviscoelas_3D.cu:772-774,
 * vx += rik[it]); QO_SRC_EXPLOSIVE follows the commented pattern at
 * cuviscoelas_3D.cu:775-776 extended to pxx+pyy+pzz. */
enum { QO_SRC_FX = 0, QO_SRC_FY = 1, QO_SRC_FZ = 2, QO_SRC_EXPLOSIVE = 3 };

/* Snapshot / state field ids (order of the 9 wavefields). */
enum { QO_VX = 0, QO_VY, QO_VZ, QO_PXX, QO_PYY, QO_PZZ, QO_PXY, QO_PXZ, QO_PYZ };

typedef struct {
    int nx, ny, nz;        /* total grid incl. C-PML (ntx, nty, ntz) */
    int pml;               /* C-PML thickness in cells (>= 2) */
    int nt;                /* number of time steps */
    float dx, dy, dz, dt;  /* m, s */
    float f0;              /* Hz: C-PML alpha_max = pi*f0 */
    float w0;              /* rad/s: reference angular frequency of the Q model */
    int src_kind;          /* QO_SRC_* */
    int sx, sy, sz;        /* true injection cell (the reference injects at s_iz+10) */
    int nrec;              /* receivers */
    const int *rx, *ry, *rz;
    int vti;               /* 0: isotropic (reference), 1: VTI (eps, delta used) */
    int snap_every;        /* 0: off; else snapshot when it % snap_every == 0 */
    int snap_field;        /* QO_VX ... QO_PYZ */
    int nthreads;          /* OpenMP threads (<=0: library default) */
    /* opt-in k-space factor of the stress operators (all 0: the reference
     * scheme): sinc^2(ks_vref |k| dt / 2) and a radial Tukey taper with
     * cut-off 2 pi ks_fcut / ks_vmax and taper ratio ks_taper */
    float ks_vref, ks_fcut, ks_vmax, ks_taper;
    /* Planted mistakes (QO_SLIP_* bits; 0 = the reference scheme).  Only for
     * tests that show a physics check fails on a plausible misreading of the
     * reference; never set otherwise. */
    int slip;
    /* Extra point sources (the receiver-side injection of a reverse-time
     * migration): after the shot's source, v_c[cell r] += inj_c[it][r] for
     * each component whose trace array [nt][ninj] is non-NULL.  Like the
     * shot's source they enter the next step's V - V_prev. */
    int ninj;
    const int *inj_x, *inj_y, *inj_z;
    const float *inj_vx, *inj_vy, *inj_vz;
    /* Wavefield history: after the step's injections, field hist_field
     * (QO_VX..QO_PYZ, or QO_PRESSURE = (pxx + pyy + pzz) / 3) of every cell
     * at every step into hist [nt][N] (float or double as the build);
     * hist_field < 0: off. */
    int hist_field;
    void *hist;
    /* Propagation medium: 0 constant-Q (the scheme), 1 dispersion only (tau
     * terms dropped), 2 lossless (gamma = 0), 3 Q-compensating (tau terms
     * negated). */
    int media;
} qo_config;

enum { QO_PRESSURE = 9 };

/* Planted mistakes for the slip tests. */
enum {
    QO_SLIP_ETA_SWAP = 1,    /* eta1 <-> eta3: |k|^-1 and |k| terms exchanged */
    QO_SLIP_NO_TAU = 2,      /* tau1 = tau2 = 0: the d/dt loss terms dropped */
    QO_SLIP_SHEAR_DPLUS = 4, /* dz vx (for pxz) with D+ instead of D- */
    QO_SLIP_PML_NODE_V = 8,  /* phi_pxx_x with node instead of half-cell a, b */
    QO_SLIP_PML_NODE_S = 16, /* phi_vxz with node instead of half-cell a, b */
    QO_SLIP_ETA_SIGN = 32    /* eta1, eta3 negated: the dispersion term reversed */
};

/* Full forward run.  Model arrays are FP32 over the total grid, index
 * ix*ny*nz + iy*nz + iz; gp/gs are gamma = atan(1/Q)/pi (caller computes, as
 * in viscoelas_3D.cpp:285-291).  eps/delta may be NULL when vti == 0.
 * Outputs (any may be NULL): rec_* [nt][nrec]; snaps [nsnap][interior window];
 * state [9][N] = final wavefields in QO_VX.. order.
 * Returns 0 on success, <0 on bad arguments. */
int qo_run_f32(const qo_config *c, const float *rik,
               const float *vp, const float *gp, const float *vs, const float *gs,
               const float *rho, const float *eps, const float *delta,
               float *rec_vx, float *rec_vy, float *rec_vz,
               float *snaps, float *state);
int qo_run_f64(const qo_config *c, const float *rik,
               const float *vp, const float *gp, const float *vs, const float *gs,
               const float *rho, const float *eps, const float *delta,
               double *rec_vx, double *rec_vy, double *rec_vz,
               double *snaps, double *state);

/* Unnormalised complex FFTs (sign -1 forward, +1 inverse), interleaved re/im.
 * 1-D: `howmany` contiguous transforms of length n.  3-D: row-major n0 x n1 x n2. */
int qo_fft1d_f32(int n, int sign, int howmany, float *data);
int qo_fft1d_f64(int n, int sign, int howmany, double *data);
int qo_fft3d_f32(int n0, int n1, int n2, int sign, float *data);
int qo_fft3d_f64(int n0, int n1, int n2, int sign, double *data);

/* Setup tables exactly as the run uses them (for closed-form tests).
 * kfold: folded |K| per axis [n]; dplus/dminus: complex shift-derivative
 * symbols along one axis [2n] (cuviscoelas_3D.cu:44-149). */
int qo_axis_tables_f64(int n, float d, double *kfold, double *dplus, double *dminus);
int qo_axis_tables_f32(int n, float d, float *kfold, float *dplus, float *dminus);
/* C-PML 1-D profiles a[n], b[n] (cuviscoelas_3D.cu:213-309). */
int qo_pml_profile_f64(int n, int pml, float d, float f0, double *a, double *b);
int qo_pml_profile_f32(int n, int pml, float d, float f0, float *a, float *b);
/* Q-model coefficients for one cell: out[5] = eta1, eta2, eta3, tau1, tau2
 * (cuviscoelas_3D.cu:195-205), FP64. */
void qo_q_coeffs(double rho, double v, double gamma, double w0, double *out);

/* The k-space factor F(|k|) for given |k| values (closed-form tests). */
int qo_kspace_factor_f64(int n, const double *kabs, float dt, float vref, float fcut, float vmax, float taper,
                         double *out);
int qo_kspace_factor_f32(int n, const float *kabs, float dt, float vref, float fcut, float vmax, float taper,
                         float *out);

int qo_version(void);

#ifdef __cplusplus
}
#endif
#endif
