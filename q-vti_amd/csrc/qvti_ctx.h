// qvti_ctx.h -- the engine context (one grid, up to S shots on one device):
// device buffers, launch arguments, run state.  Shared by the translation
// units of the host side (qvti_engine.hip, qvti_rtm.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <vector>

#include "qvti.h"
#include "qvti_aux.h"
#include "qvti_launch.h"
#include "qvti_pencils.h"

namespace qvti {
enum KClass {
    KC_TICK = 0, KC_VEL_X, KC_VEL_Y, KC_VEL_Z, KC_Y_FWD, KC_XPASS, KC_Y_INV, KC_Z_INV,
    KC_STR_X, KC_STR_Y, KC_STR_Z, KC_RECORD, KC_SNAP, KC_PROBE, KC_COUNT
};
struct AxisDev {
    int n;
    float d;
    float2 *tw = nullptr, *dp = nullptr, *dm = nullptr;
    float *kf = nullptr, *a = nullptr, *b = nullptr, *ah = nullptr, *bh = nullptr;
};
}  // namespace qvti

using namespace qvti;

struct qvti_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    qvti_grid g{};
    long long N = 0, nrows = 0;
    long long Nb = 0;  // blk_len(N): per-shot field length of the blocked groups
    int S = 1;  // max shots
    AxisDev ax[3];
    const LenOps* ops[3] = {nullptr, nullptr, nullptr};  // kernels of the x, y, z lengths
    int split[3] = {0, 0, 0};                            // their split variants (len_ops)
    // model
    float* coef = nullptr;  // blocked coefficient groups G1 (15 fields) then G2 (10, VTI)
    float* buoy = nullptr;  // [3][N]
    bool has_model = false;
    int media = 0;  // QVTI_MEDIA_*: applied by the next qvti_set_model
    // state, field-major [field][shot][N]
    float *v = nullptr, *sig = nullptr, *dv = nullptr, *H = nullptr, *bufA = nullptr, *bufB = nullptr;
    float *phiv[3] = {nullptr, nullptr, nullptr}, *phis[3] = {nullptr, nullptr, nullptr};
    long long slab[3] = {0, 0, 0};
    int nphis[3] = {5, 5, 4};
    // run
    int nt = 0, nshots = 0, steps_done = 0, begun = 0;
    float* d_rik = nullptr;
    int* d_it = nullptr;
    int* d_probe = nullptr;
    int* d_src_kind = nullptr;
    long long* d_src_ip = nullptr;
    long long* d_rec_ip = nullptr;
    int* d_rec_off = nullptr;
    int total_rec = 0;
    float* d_rec = nullptr;
    long long rec_cap = 0;
    int probe_every = 50;
    float probe_limit = 1e30f;
    // snapshot of the box [snap_lo, snap_hi) (default: the interior window)
    int snap_field = 0, snap_every = 0;
    int snap_lo[3] = {0, 0, 0}, snap_hi[3] = {0, 0, 0};
    float* snap_host = nullptr;
    long long snap_cap = 0;
    float* d_snap = nullptr;
    long long d_snap_alloc = 0;  // floats allocated at d_snap (only grows)
    long long snap_lim = 0;      // floats this run may write: min(box * snapshots, snap_cap)
    // graphs
    bool use_graphs = true;
    hipGraphExec_t gexec = nullptr;
    // timing
    bool timing = false;
    int ev_max = 0, ev_used = 0;
    std::vector<hipEvent_t> ev;
    std::vector<int> ev_class;
    long long kc_launches[KC_COUNT] = {};
    double kc_ms[KC_COUNT] = {};
    // launch configuration
    StridedArgs a_vx{}, a_vy{}, a_sx{}, a_sy{};
    VelZArgs a_vz{};
    ZRowsArgs a_zi{};
    CStridedArgs a_yf{}, a_yi{};
    XPassArgs a_xp{};
    StressZArgs a_sz{};
    RecArgs a_rec{};
    dim3 g_vx, g_vy, g_vz, g_zr, g_cy, g_xp, g_sz, g_rec, g_sx, g_sy, g_yi;
    int P[3] = {0, 0, 0};
    // tile width per kernel class (strided-axis kernels), chosen by the autotuner
    int pw[KC_COUNT] = {};
    // fields per workgroup of the strided kernels: grid z = 1 (every field in
    // one workgroup) or the field count (one field per workgroup)
    int jz[KC_COUNT] = {};
    // XCD placement: consecutive row tiles (vel_z, str_z) and consecutive
    // planes (strided kernels) that run on one XCD (tile_shot, plane_tile)
    int run_rows = 1, run_planes = 1;
    bool autotune = true;
    int tuned = 0;  // shot count the tile widths were tuned for (0: not tuned)
    // opt-in k-space correction / wavenumber filter (off: the reference scheme)
    bool ks = false;
    float ks_vref = 0.f, ks_fcut = 0.f, ks_vmax = 0.f, ks_taper = 0.f;
    float* H2 = nullptr;  // [A2 v x,y,z][A2 dv x,y,z], allocated on first use
    long long bytes = 0;
    std::vector<void*> allocs;
};


// Field k of a field-major [field][shot][N] state buffer, shot 0.
inline float* fld(float* base, int k, const qvti_ctx* c) { return base + (long long)k * c->S * c->N; }
