/*
 * qvti.h -- C ABI of the MI355X engine for the Q-VTI time-stepping loop.
 *
 * The boundary replaces the reference's C ABI (Myfunctions.h:1-113).  Plain
 * pointers and sizes only; every entry point returns an int status (0 = ok,
 * negative = error, message in qvti_last_error()).  Model arrays are FP32 over
 * the TOTAL grid (C-PML included), index ix*ny*nz + iy*nz + iz, as in the
 * reference (Myfunctions.h:83, viscoelas_3D.cpp:576-655).  Records are FP32
 * [nt][R] per component with the receiver index fastest
 * (cuviscoelas_3D.cu:789-791).
 *
 * One context per host thread; a context owns one HIP stream.  The serving
 * queue (qvti_server_*) is internally synchronised.
 */
#ifndef QVTI_H
#define QVTI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes */
#define QVTI_OK 0
#define QVTI_E_ARG (-1)         /* invalid argument / shape */
#define QVTI_E_DEVICE (-2)      /* HIP runtime error or no device */
#define QVTI_E_NOMEM (-3)       /* device allocation failed */
#define QVTI_E_UNSUPPORTED (-4) /* axis length not an even 2/3/5-smooth number in [16, 512] */
#define QVTI_E_STATE (-5)       /* call out of order (e.g. no model set) */
#define QVTI_E_UNSTABLE (-6)    /* blow-up probe fired (NaN/Inf or |v| > limit) */
#define QVTI_E_BUDGET (-7)      /* request exceeds the server's HBM budget */
#define QVTI_E_CANCELLED (-8)   /* request cancelled */
#define QVTI_E_PENDING (-9)     /* request not finished yet (qvti_server_poll) */

/* ---- source kinds.  QVTI_SRC_FX is the reference default: vx += rik[it]
 * after the stress update (cuviscoelas_3D.cu:772-774).  QVTI_SRC_EXPLOSIVE
 * adds rik[it] to pxx, pyy and pzz (the commented pattern at
 * cuviscoelas_3D.cu:775-776). */
enum { QVTI_SRC_FX = 0, QVTI_SRC_FY = 1, QVTI_SRC_FZ = 2, QVTI_SRC_EXPLOSIVE = 3 };

/* ---- wavefield ids (snapshots, state readback) */
enum { QVTI_VX = 0, QVTI_VY, QVTI_VZ, QVTI_PXX, QVTI_PYY, QVTI_PZZ, QVTI_PXY, QVTI_PXZ, QVTI_PYZ };

typedef struct qvti_ctx qvti_ctx;

/* Grid and scheme constants; replaces the scalar arguments of
 * cuda_Device_malloc / cuda_forward_acoustic_3D (Myfunctions.h:77-96). */
typedef struct {
    int nx, ny, nz;   /* total grid (ntx, nty, ntz), each even, 2/3/5-smooth, 16..512 */
    int pml;          /* C-PML thickness (>= 2, 2*pml <= each n) */
    float dx, dy, dz; /* m */
    float dt;         /* s */
    float f0;         /* Hz: C-PML alpha_max = pi*f0 (cuviscoelas_3D.cu:238) */
    float w0;         /* rad/s: Q-model reference frequency (viscoelas_3D.cpp:109) */
    int vti;          /* 0 isotropic (reference), 1 VTI (Thomsen eps, delta) */
} qvti_grid;

/* One shot's geometry.  The injection cell is the TRUE cell: the reference
 * injects at (s_ix, s_iy, s_iz+10) (cuviscoelas_3D.cu:772); the
 * reference-compatible entry point applies that +10 itself. */
typedef struct {
    int kind;          /* QVTI_SRC_* */
    int sx, sy, sz;    /* injection cell */
    int nrec;          /* receivers R */
    const int *rx, *ry, *rz; /* receiver cells [R] */
} qvti_shot;

/* Layout-identical to the reference's struct Source (Myfunctions.h:6-10). */
typedef struct {
    int s_iz, s_ix, s_iy, r_iz;
    int *r_ix, *r_iy;
    int r_n;
} qvti_ref_source;

/* ---- lifecycle.  Replaces getdevice (Myfunctions.h:3-4). */
int qvti_device_count(int *n);

/* Allocate a context for up to max_shots concurrent shots of one grid on
 * `device`.  Replaces cuda_Device_malloc + cuda_Host_initialization
 * (Myfunctions.h:88-96, cuviscoelas_3D.cu:1658-1916). */
int qvti_create(qvti_ctx **out, const qvti_grid *g, int max_shots, int device);

/* Upload a model and build the FP64-then-FP32 coefficient fields, C-PML
 * profiles and spectral tables.  gama_p/gama_s are gamma = atan(1/Q)/pi
 * computed by the caller (viscoelas_3D.cpp:285-291).  eps/delta are the
 * Thomsen fields (required iff grid.vti).  Replaces the H2D copies and setup
 * kernels of cuda_forward_acoustic_3D (cuviscoelas_3D.cu:1181-1249). */
int qvti_set_model(qvti_ctx *c, const float *velp, const float *gama_p, const float *vels,
                   const float *gama_s, const float *rho, const float *eps, const float *delta);

/* One shot, synchronous: the per-shot loop of cuda_forward_acoustic_3D
 * (cuviscoelas_3D.cu:1137-1652).  rik [nt]; rec_* [nt][R] (any may be NULL). */
int qvti_forward(qvti_ctx *c, int nt, const float *rik, const qvti_shot *shot,
                 float *rec_vx, float *rec_vy, float *rec_vz);

/* nshots shots of the same model stepped together (batched launches, one HIP
 * graph per step chunk).  rec_* [nshots][nt][R_s] with R_s each shot's nrec
 * packed back to back.  No counterpart in the reference (its GPU_N loop runs
 * one shot per GPU, cuviscoelas_3D.cu:1257-1629). */
int qvti_forward_batch(qvti_ctx *c, int nt, const float *rik, const qvti_shot *shots, int nshots,
                       float *rec_vx, float *rec_vy, float *rec_vz);

/* Shots per launch for nshots shots on this context: tuned by timing silent
 * runs of widths 1, 2, 4, ... up to min(max_shots, nshots) on the context
 * (which ends any run in progress), once per grid, k-space mode and
 * min(capacity, shot count), and kept for the process.  Results do not depend
 * on the width.  Writes the width; the widths tried and their ms per
 * shot-step go to trial_width / trial_ms (up to max_trials entries, either
 * may be NULL).  Returns the number of widths tried.  No counterpart in the
 * reference (one shot per GPU, viscoelas_3D.cpp:244-256). */
int qvti_batch_width(qvti_ctx *c, const qvti_shot *shots, int nshots, int *width, int max_trials,
                     int *trial_width, double *trial_ms);
/* Any number of shots of the current model, run in groups of the tuned
 * width; records as qvti_forward_batch ([shot][nt][R_s] back to back). */
int qvti_forward_shots(qvti_ctx *c, int nt, const float *rik, const qvti_shot *shots, int nshots,
                       float *rec_vx, float *rec_vy, float *rec_vz);

/* Same argument list and meaning as the reference's cuda_forward_acoustic_3D
 * (Myfunctions.h:77-85) minus the MultiGPU plan: uploads the model, runs shot
 * ss[is] with the reference's x-force at (s_ix, s_iy, s_iz+10) and receivers
 * (r_ix[ir], r_iy[ir], r_iz), ir < rnx_max*rny_max, and writes the three
 * records [nt][rnx_max*rny_max].  velp_max is ignored (the C-PML uses 5000,
 * cuviscoelas_3D.cu:235); myid, nx, ny, nz, dr are accepted and unused, as in
 * the reference. */
int qvti_forward_acoustic_3D(qvti_ctx *c, int myid, int is, int nt, int ntx, int nty, int ntz, int ntp,
                             int nx, int ny, int nz, int pml, float dx, float dy, float dz, float dt,
                             float f0, float w0, float velp_max, const float *rik, const float *velp,
                             const float *gama_p, const float *vels, const float *gama_s,
                             const float *rho, const qvti_ref_source *ss, int rnmax, int rnx_max,
                             int rny_max, int dr, float *record, float *record2, float *record3);

/* Snapshot hook: every `every` steps (it % every == 0, after the step's
 * source injection) copy the snapshot window (default: the interior
 * [pml, n-pml)^3) of `field`, ix slowest, into host_buf (capacity in
 * floats).  every = 0 disables.
 * The reference's disabled dump (cuviscoelas_3D.cu:1615-1627) used 200. */
int qvti_set_snapshot(qvti_ctx *c, int field, int every, float *host_buf, long long capacity);
/* Snapshot window: the box [lo, hi) of total-grid indices (x, y, z) that the
 * snapshots copy, ix slowest.  Default: the interior [pml, n - pml) on every
 * axis, as the reference's dump (cuviscoelas_3D.cu:1615-1627). */
int qvti_set_snapshot_window(qvti_ctx *c, const int *lo, const int *hi);

/* ---- propagation medium, the reference's media_type flag
 * (viscoelas_3D.cpp:57-60): QVTI_MEDIA_VISCO the constant-Q scheme (default);
 * DISPERSION its phase dispersion without the amplitude loss (tau terms
 * dropped); LOSSLESS gamma = 0; QCOMP Q-compensating: the amplitude loss
 * turned into a gain of the same size, the dispersion kept, for back-
 * propagating recorded data (pair it with the wavenumber filter of
 * qvti_set_kspace, which bounds the gain).  Applies from the next
 * qvti_set_model. */
enum { QVTI_MEDIA_VISCO = 0, QVTI_MEDIA_DISPERSION = 1, QVTI_MEDIA_LOSSLESS = 2, QVTI_MEDIA_QCOMP = 3 };
int qvti_set_media(qvti_ctx *c, int media);

/* ---- reverse-time migration of one shot.  The reference declares the flags
 * and never builds the mode (RTM_type, imaging_condition, reconstruction,
 * viscoelas_3D.cpp:42-65; checkpoints every check_steps = 500 steps,
 * viscoelas_3D.cpp:223-229).  Imaging conditions use the reference's codes. */
enum { QVTI_IC_EXA = 1, QVTI_IC_SRCNORM = 3, QVTI_IC_XCORR = 4 };
enum { QVTI_FIELD_PRESSURE = 9 }; /* (pxx + pyy + pzz) / 3 */
typedef struct {
    int condition;        /* QVTI_IC_*: excitation amplitude R(t_ex)/S(t_ex);
                           * source-normalised sum S R / (sum S^2 + eps max sum S^2);
                           * cross-correlation sum S R */
    int field;            /* imaged wavefield: QVTI_VX..QVTI_PYZ or QVTI_FIELD_PRESSURE */
    int checkpoint_every; /* source-wavefield checkpoint interval C in steps: the
                           * source run is recomputed in segments of C steps;
                           * device memory 2 C nx ny nz floats (not for EXA) */
    float eps;            /* stabilisation: source-normalised adds eps max sum S^2 to
                           * the denominator; EXA images only where
                           * |S(t_ex)| > eps max |S(t_ex)| */
} qvti_rtm_opts;
/* src and rcv: two contexts of one grid (C-PML, dt, device), each with a
 * model set: src propagates the shot's source (rik [nt]); rcv back-propagates
 * the records rec_* [nt][R] of the shot's R receivers (any may be NULL),
 * injected time-reversed as forces at the receiver cells (its medium may be
 * Q-compensating).  Image [nx][ny][nz]; illum (may be NULL): sum S^2, or for
 * EXA the excitation amplitude S(t_ex).  Both contexts' runs end (begin them
 * again before stepping).  QVTI_E_UNSTABLE if either wavefield blew up. */
int qvti_rtm(qvti_ctx *src, qvti_ctx *rcv, int nt, const float *rik, const qvti_shot *shot,
             const float *rec_vx, const float *rec_vy, const float *rec_vz, const qvti_rtm_opts *opts,
             float *image, float *illum);

/* Release everything.  Replaces cuda_Device_free (Myfunctions.h:98-102). */
int qvti_destroy(qvti_ctx *c);

/* Thread-local message of the last failing call on this thread. */
const char *qvti_last_error(void);

/* ---- stepping API (benchmarks, tests, serving) */
int qvti_begin(qvti_ctx *c, int nt, const float *rik, const qvti_shot *shots, int nshots);
int qvti_step(qvti_ctx *c, int nsteps);  /* enqueue asynchronously */
int qvti_sync(qvti_ctx *c);              /* wait; returns QVTI_E_UNSTABLE if the probe fired */
int qvti_finish(qvti_ctx *c, float *rec_vx, float *rec_vy, float *rec_vz);
int qvti_get_state(qvti_ctx *c, int shot, int field, float *out); /* [nx][ny][nz] */
int qvti_steps_done(qvti_ctx *c);
/* Checkpoint / resume of a run in progress (every shot of the context).  The
 * reference allocates checkpoint bookkeeping it never uses (check_steps,
 * N_cp, t_cp at viscoelas_3D.cpp:223-229, 457).  The image holds the state
 * that carries from step to step - velocities, stresses, C-PML memory - plus
 * the step count, the records gathered so far and the snapshots taken so far.
 * qvti_checkpoint waits for the queued steps.  qvti_restore loads an image
 * into a context of the same grid that was given the same qvti_begin (nt,
 * wavelet, shots, snapshot settings) and has not stepped yet; the run then
 * continues from the saved step, bit-identical to an uninterrupted run. */
long long qvti_checkpoint_bytes(qvti_ctx *c);
int qvti_checkpoint(qvti_ctx *c, void *buf, long long capacity);
int qvti_restore(qvti_ctx *c, const void *buf, long long size);
/* Per-kernel-class HIP event timing of the launches enqueued while enabled.
 * qvti_kernel_times returns, per class, launches and total milliseconds. */
int qvti_enable_kernel_timing(qvti_ctx *c, int on);
int qvti_kernel_times(qvti_ctx *c, int max_classes, char *names, int name_len, long long *launches,
                      double *total_ms);
/* Algorithmic HBM bytes one launch of each kernel class moves (for roofline). */
int qvti_kernel_bytes(qvti_ctx *c, int max_classes, double *bytes);
/* Use HIP graphs for chunks of steps (default on). */
int qvti_set_graphs(qvti_ctx *c, int on);
/* Tile width of the strided-axis kernels.  mode 1 (default): autotune - at
 * the first qvti_begin for a shot count, every available width of each
 * strided kernel class is timed on this context and the fastest kept.
 * mode 0: the narrowest tiles; mode 2 or 4: that width (or the widest
 * available below it) for every class; + 16: one field per workgroup.
 * Results are identical for every choice: a pencil's arithmetic does not
 * depend on the tile. */
int qvti_set_autotune(qvti_ctx *c, int mode);
/* Pencils per tile and workgroups per tile (1: all fields of the launch in
 * one workgroup; else one field per workgroup) chosen for each kernel class
 * (0 for row kernels), in the class order of qvti_kernel_times; split may be
 * NULL.  Returns the number of classes. */
int qvti_tile_choice(qvti_ctx *c, int max_classes, int *pencils, int *split);
/* Opt-in k-space time-stepping correction and wavenumber filter of the
 * stress operators.  Both are off by default (the reference scheme); the
 * reference has them switched off: sinc2nd = 1 at cuviscoelas_3D.cu:183-190,
 * the filter array commented out at Myfunctions.h:69 and its parameters
 * unused at viscoelas_3D.cpp:111-112.
 *  vref > 0: every spectral operator of the stress update (A1 = |k|^-1,
 *    A2 = 1, A3 = |k|) is multiplied by sinc^2(vref |k| dt / 2),
 *    sinc(x) = sin(x)/x, the reference's disabled sinc2nd with
 *    velp_max = vref.  Waves of speed vref then propagate without time
 *    dispersion, and a homogeneous medium with speeds <= vref is stable at
 *    any dt.
 *  fcut > 0: the same operators are multiplied by a radial Tukey taper of
 *    |k|: 1 up to (1 - taper_ratio) kc, a half cosine down to 0 at kc, 0
 *    beyond, with kc = 2 pi fcut / vmax (fcut in Hz, vmax in m/s,
 *    taper_ratio in [0, 1]).
 *  vref = 0 and fcut = 0: both off.
 * Takes effect at the next qvti_begin; a run in progress must be begun
 * again, as after qvti_set_snapshot.  The first call that switches it on allocates 6 more fields per
 * shot; each step then runs three more complex 3-D inverse transforms. */
int qvti_set_kspace(qvti_ctx *c, float vref, float fcut, float vmax, float taper_ratio);
/* Two-level FFT split of each axis (0: the default N1 x N2 of the length;
 * 1: an alternative for lengths 64 96 100 128 160 200 256 400).  The split
 * changes the FP32 rounding (not the scheme), so it is never picked by a
 * run's own timing: qvti_set_split fixes it for a context; qvti_tune_split
 * times both variants per axis on the context (silent runs, ending any run in
 * progress), applies the faster and stores the choice for the grid (nx, ny,
 * nz, vti) in the process table that every later qvti_create of that grid
 * uses.  qvti_save_wisdom / qvti_load_wisdom write / merge that table as text
 * ("qvti-wisdom 1", then "nx ny nz vti split_x split_y split_z" lines).
 * ms_per_step (may be NULL) receives [axis][variant] timings (6 values; the
 * entries of axes without an alternative are left unset).  No counterpart in
 * the reference (cuFFT plans its own factorisation, cuviscoelas_3D.cu:1703). */
int qvti_set_split(qvti_ctx *c, int axis, int variant);
int qvti_get_split(qvti_ctx *c, int *variants);
int qvti_tune_split(qvti_ctx *c, int *variants, double *ms_per_step);
int qvti_save_wisdom(const char *path);
int qvti_load_wisdom(const char *path);
/* Blow-up probe interval in steps (0 disables; default 50) and limit on |v|. */
int qvti_set_probe(qvti_ctx *c, int every, float limit);
/* Device bytes this context holds. */
long long qvti_device_bytes(qvti_ctx *c);
/* Device bytes a context for grid g with max_shots would hold. */
long long qvti_estimate_bytes(const qvti_grid *g, int max_shots);

/* ---- standalone transforms on the engine's kernels (tests / tuning).
 * In-place unnormalised 3-D C2C of interleaved complex data [nx][ny][nz]. */
int qvti_fft3d(int nx, int ny, int nz, int sign, float *host_data);
/* Allocation-free variant of qvti_fft3d for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  Same transform,
 * but on DEVICE data d_data (interleaved complex [nx][ny][nz], caller-owned),
 * enqueued on `stream` (a hipStream_t; NULL: the null stream) without
 * waiting.  Its only other device memory is a static 12 KiB per-device
 * scratch array holding the three axes' twiddle tables (each axis <= 512;
 * longer axes are rejected with QVTI_E_UNSUPPORTED); every call rebuilds the
 * tables it uses.  Calls that overlap on one device race on that array:
 * serialise them (one stream, or events).  Calls no hipMalloc/hipFree. */
int qvti_fft3d_noalloc(int nx, int ny, int nz, int sign, float *d_data, void *stream);
/* 1-D shift-derivative of a real field along `axis` (0 x, 1 y, 2 z) with the
 * reference's D+ (sym 0) or D- (sym 1) symbol, /n normalised. */
int qvti_deriv1d(int nx, int ny, int nz, float d, int axis, int sym, const float *in, float *out);

/* ---- host helpers mirroring the reference host (viscoelas_3D.cpp) */
/* Ricker wavelet, viscoelas_3D.cpp:494-502. */
int qvti_ricker(int nt, float dt, float f0, float *rik);
/* gamma = atan(1/Q)/pi, viscoelas_3D.cpp:287-288. */
int qvti_gamma_from_q(const float *q, float *gamma, long long n);
/* Receiver grid of the reference: r_ix = pml + ir/rny*dr, r_iy = pml + (ir%rny)*dr
 * (viscoelas_3D.cpp:181-185). */
int qvti_reference_receivers(int pml, int rnx_max, int rny_max, int dr, int *r_ix, int *r_iy);

/* ---- data formats on either side of the path (host only) */
/* Edge-replicate an interior model [(ntx-2pml)][(nty-2pml)][(ntz-2pml)] into
 * the total grid [ntx][nty][ntz]: the reference's padding of its model files
 * into the C-PML frame (viscoelas_3D.cpp:576-655). */
int qvti_pad_model(const float *interior, int ntx, int nty, int ntz, int pml, float *total);
/* Read an interior model from a raw FP32 file (ix slowest, the reference's
 * input format, viscoelas_3D.cpp:541-546) and pad it into total. */
int qvti_load_model_raw(const char *path, int ntx, int nty, int ntz, int pml, float *total);
/* Write n FP32 values raw: record gathers [nt][R] as the reference dumps them
 * (viscoelas_3D.cpp:416-434) or interior snapshots (cuviscoelas_3D.cu:1615-1627). */
int qvti_write_raw(const char *path, const float *data, long long n);

/* Seeded synthetic models for the benchmark configurations (splitmix64-seeded
 * xoshiro256**, u = (x >> 11) * 2^-53).  Outputs FP32 [nx][ny][nz]; eps/delta
 * may be NULL.  Q is returned through qp/qs (gamma via qvti_gamma_from_q).
 *   kind 0: homogeneous; p = {vp, vs, rho, Qp, Qs}
 *   kind 1: nlay equal flat layers in z; per layer vp ~ U[p0,p1], eps ~ U[0,p2],
 *           delta ~ U[0,eps], Qp ~ U[p3,p4]; vs = vp/1.73, rho = 310 vp^0.25
 *           (Gardner), Qs = Qp/2
 *   kind 2: white N(0,1) noise, Gaussian filter sigma = p5 cells, min-max to
 *           t in [0,1]: vp = p0 + (p1-p0) t, Qp = p3 + (p4-p3) t,
 *           eps = p6 + (p2-p6)(1-t), delta = eps/2, vs = vp/1.73,
 *           rho = 310 vp^0.25, Qs = Qp/2 */
int qvti_make_model(int kind, uint64_t seed, int nx, int ny, int nz, int nlay, const double *p,
                    float *vp, float *vs, float *rho, float *qp, float *qs, float *eps, float *delta);

/* ---- serving: continuous batching of shot requests of uneven size */
typedef struct qvti_server qvti_server;
typedef struct {
    qvti_grid grid;
    int nt;
    const float *velp, *gama_p, *vels, *gama_s, *rho, *eps, *delta; /* copied at submit */
    const float *rik;                                               /* [nt], copied */
    qvti_shot shot;                                                 /* receivers copied */
} qvti_request;
typedef struct {
    int status;          /* QVTI_OK, QVTI_E_* */
    double t_submit_s, t_admit_s, t_done_s; /* server clock */
    int nt, nrec;
} qvti_result_info;

int qvti_server_create(qvti_server **out, int device, long long hbm_budget_bytes, int max_active,
                       int steps_per_slice);
int qvti_server_submit(qvti_server *s, const qvti_request *r, uint64_t *id);
/* QVTI_E_PENDING while running; otherwise the request's final status.  rec_*
 * [nt][nrec] copied out when finished (may be NULL). */
int qvti_server_poll(qvti_server *s, uint64_t id, qvti_result_info *info, float *rec_vx,
                     float *rec_vy, float *rec_vz);
int qvti_server_wait(qvti_server *s, uint64_t id, double timeout_s);
int qvti_server_cancel(qvti_server *s, uint64_t id);
double qvti_server_clock(qvti_server *s);
int qvti_server_destroy(qvti_server *s);

#ifdef __cplusplus
}
#endif
#endif
