// qvti_engine.hip -- host side of the MI355X engine: context, FP64-built
// tables, launch sequence, HIP-graph stepping and the C ABI of include/qvti.h.
//
// Reference being replaced: the per-shot driver cuda_forward_acoustic_3D
// (cuviscoelas_3D.cu:1137-1652) with its allocation/free helpers
// (cuviscoelas_3D.cu:1655-2148) and setup kernels (cuviscoelas_3D.cu:22-410).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "qvti.h"
#include "qvti_aux.h"
#include "qvti_ctx.h"
#include "qvti_internal.h"
#include "qvti_launch.h"
#include "qvti_pencils.h"

using namespace qvti;

// ------------------------------------------------------------ errors -----
// qvti_fail / qvti_last_error live in qvti_error.cpp.

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            (void)hipGetLastError(); /* reported here, not again by a later check */       \
            return qvti_fail(e_ == hipErrorOutOfMemory ? QVTI_E_NOMEM : QVTI_E_DEVICE,     \
                             "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, \
                             __LINE__);                                                    \
        }                                                                                  \
    } while (0)

// ------------------------------------------------------------ lengths ----
constexpr int kMaxN = 512;

// The launch records of every supported axis length (qvti_lengths.h); each
// is defined in its build group of qvti_kernels.hip.
const LenOps* qvti::len_ops(int n, int variant)
{
#define QVTI_LEN_PTR(n_, a, b, g) &kLenOps_##n_,
    static const LenOps* const all[] = {QVTI_LENGTHS(QVTI_LEN_PTR)};
#undef QVTI_LEN_PTR
#define QVTI_ALT_PTR(n_, a, b, g) &kLenOps_Alt_##n_,
    static const LenOps* const alt[] = {QVTI_ALT_LENGTHS(QVTI_ALT_PTR)};
#undef QVTI_ALT_PTR
    if (variant == 1) {
        for (const LenOps* o : alt)
            if (o->n == n) return o;
        return nullptr;
    }
    if (variant != 0) return nullptr;
    for (const LenOps* o : all)
        if (o->n == n) return o;
    return nullptr;
}

const char* qvti::supported_lengths()
{
#define QVTI_LEN_STR(n_, a, b, g) " " #n_
    static const char* s = QVTI_LENGTHS(QVTI_LEN_STR) + 1;
#undef QVTI_LEN_STR
    return s;
}

// Effective (PW, pencils per tile, threads) of length n at requested width pw.
struct TileW {
    int pw, P, T;
};
static TileW tile_w(const LenOps* o, int pw)
{
    if (!o) return TileW{0, 0, 0};
    const int i = pw_index(*o, pw);
    return TileW{1 << i, o->P[i], o->T[i]};
}

static bool supported_n(int n) { return len_ops(n) != nullptr; }

// The reference's truncated pi (cuviscoelas_3D.cu:15).
static const double kRefPi = 3.1415926;

// D+ / D- symbols along one axis, cuviscoelas_3D.cu:44-149, with the FP32
// intermediates the reference keeps (dk, tmp, K = tmp*m).  The Nyquist symbol
// keeps its real part only (the `.x` of the reference's C2C result sees the
// Hermitian part), and 1/n is folded in (the reference's `/ntp`).
static void axis_symbols(int n, float d, std::vector<float2>& dp, std::vector<float2>& dm, std::vector<float>& kf)
{
    dp.resize(n);
    dm.resize(n);
    kf.resize(n);
    const float dk = (float)(1.0 / n / d);
    const float tmp = (float)(2 * kRefPi * dk);
    const double inv = 1.0 / n;
    for (int i = 0; i < n; ++i) {
        const int m = i <= n / 2 ? i : n - i;
        kf[i] = (float)(2 * kRefPi / n / d * m);
        const float K = tmp * (float)m;
        const double s = sin(m * kRefPi / n), c = cos(m * kRefPi / n);
        float pr = (float)(-(double)K * s), mr = (float)((double)K * s);
        float pim = (float)((double)K * c), mim = (float)((double)K * c);
        if (i > n / 2) { pim = -pim; mim = -mim; }
        if (2 * i == n) { pim = 0.f; mim = 0.f; }
        dp[i] = make_float2((float)(pr * inv), (float)(pim * inv));
        dm[i] = make_float2((float)(mr * inv), (float)(mim * inv));
    }
}

// C-PML a(x), b(x), cuviscoelas_3D.cu:213-309: R = 1e-2, n1=2, n2=1, n3=2,
// vmax forced to 5000 (:235), kappa = 1; plus half-cell averages.
static void pml_profile(int n, int pml, float d, float f0, std::vector<float>& a, std::vector<float>& b,
                        std::vector<float>& ah, std::vector<float>& bh)
{
    a.resize(n);
    b.resize(n);
    ah.resize(n);
    bh.resize(n);
    const int n1 = 2, n2 = 1, n3 = 2;
    const float R = 1e-2f, vmax = 5000.0f;
    const float alpha_max = (float)(kRefPi * f0);
    const float om_max = (float)((1 + n1 + n2) * vmax * log(1.0 / R) / ((n1 + n2 - 1) * pml * d));
    for (int i = 0; i < n; ++i) {
        const int j = i < pml ? i : (i > n - 1 - pml ? n - 1 - i : -1);
        float alpha, om;
        if (j >= 0) {
            alpha = alpha_max * powf((float)(1.0 * j / (pml - 1)), (float)n3);
            om = om_max * powf((float)(1.0 * (pml - 1 - j) / pml), (float)(n1 + n2));
        } else {
            alpha = alpha_max;
            om = 0.f;
        }
        a[i] = alpha + om / 1.0f;
        b[i] = om / 1.0f;
    }
    for (int i = 0; i < n; ++i) {
        ah[i] = i + 1 < n ? 0.5f * (a[i] + a[i + 1]) : a[i];
        bh[i] = i + 1 < n ? 0.5f * (b[i] + b[i + 1]) : b[i];
    }
}


// ------------------------------------------------------------ context ----
// (struct qvti_ctx: qvti_ctx.h)
namespace {
using namespace qvti;
const char* kClassNames[KC_COUNT] = {"tick", "vel_x", "vel_y", "vel_z_fwd", "y_fwd", "xpass", "y_inv",
                                     "z_inv", "str_x", "str_y", "str_z", "record", "snapshot", "probe"};
}  // namespace

static int dalloc(qvti_ctx* c, void** p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();  // the failure is reported here, not by a later launch check
        return qvti_fail(QVTI_E_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    c->allocs.push_back(*p);
    c->bytes += (long long)bytes;
    return QVTI_OK;
}

template <class T>
static int upload(qvti_ctx* c, T** dst, const std::vector<T>& src)
{
    int rc = dalloc(c, (void**)dst, sizeof(T) * src.size());
    if (rc) return rc;
    HIPCHK(hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
    return QVTI_OK;
}

static void release(qvti_ctx* c)
{
    if (c->gexec) hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    for (auto e : c->ev) hipEventDestroy(e);
    c->ev.clear();
    for (void* p : c->allocs) hipFree(p);
    c->allocs.clear();
    if (c->stream) hipStreamDestroy(c->stream);
    c->stream = nullptr;
}

// ------------------------------------------------------- sizes / memory --
static long long state_fields() { return 3 + 6 + 3 + 9 + 15 + 15; }

extern "C" long long qvti_estimate_bytes(const qvti_grid* g, int max_shots)
{
    if (!g || max_shots < 1) return -1;
    const long long N = (long long)g->nx * g->ny * g->nz;
    const long long p2 = 2LL * g->pml;
    const long long slabs = 8LL * (p2 * g->ny * g->nz + (long long)g->nx * p2 * g->nz) + 7LL * g->nx * g->ny * p2;
    const long long model = (g->vti ? 20 : 10) + 5 + 3;
    return 4LL * (N * (state_fields() * max_shots + model) + slabs * max_shots) + (64LL << 20);
}

extern "C" int qvti_device_count(int* n)
{
    if (!n) return qvti_fail(QVTI_E_ARG, "n is NULL");
    int k = 0;
    hipError_t e = hipGetDeviceCount(&k);
    if (e != hipSuccess) {
        *n = 0;
        return qvti_fail(QVTI_E_DEVICE, "hipGetDeviceCount: %s", hipGetErrorString(e));
    }
    *n = k;
    return QVTI_OK;
}

extern "C" long long qvti_device_bytes(qvti_ctx* c) { return c ? c->bytes : -1; }

hipStream_t qvti_ctx_stream(qvti_ctx* c) { return c ? c->stream : nullptr; }

static int build_launch_config(qvti_ctx* c);
static void apply_tiles(qvti_ctx* c);

// ------------------------------------------------------- split wisdom ---
// Split variant per axis chosen for a grid (nx, ny, nz, vti), kept for the
// process and written to / read from a wisdom file.  The FP32 rounding of a
// run depends on the splits, so a context takes them from this table - a
// stored choice, not a timing of its own.
namespace {
using GridKey = std::array<int, 4>;
std::mutex g_split_mu;
std::map<GridKey, std::array<int, 3>> g_split;
GridKey grid_key(const qvti_grid& g) { return GridKey{g.nx, g.ny, g.nz, g.vti ? 1 : 0}; }
}  // namespace

static std::array<int, 3> stored_split(const qvti_grid& g)
{
    std::lock_guard<std::mutex> lk(g_split_mu);
    auto it = g_split.find(grid_key(g));
    return it == g_split.end() ? std::array<int, 3>{0, 0, 0} : it->second;
}

extern "C" int qvti_create(qvti_ctx** out, const qvti_grid* g, int max_shots, int device)
{
    if (!out || !g) return qvti_fail(QVTI_E_ARG, "NULL argument");
    *out = nullptr;
    if (max_shots < 1 || max_shots > 4096) return qvti_fail(QVTI_E_ARG, "max_shots %d out of range", max_shots);
    const int dims[3] = {g->nx, g->ny, g->nz};
    for (int a = 0; a < 3; ++a)
        if (!supported_n(dims[a]))
            return qvti_fail(QVTI_E_UNSUPPORTED,
                             "axis %d length %d: supported axis lengths are %s", a, dims[a], supported_lengths());
    if (g->pml < 2 || 2 * g->pml > g->nx || 2 * g->pml > g->ny || 2 * g->pml > g->nz)
        return qvti_fail(QVTI_E_ARG, "pml %d invalid for grid %dx%dx%d", g->pml, g->nx, g->ny, g->nz);
    if (!(g->dx > 0 && g->dy > 0 && g->dz > 0 && g->dt > 0 && g->w0 > 0 && g->f0 > 0))
        return qvti_fail(QVTI_E_ARG, "dx, dy, dz, dt, f0, w0 must be positive");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return qvti_fail(QVTI_E_DEVICE, "no HIP device");
    if (device < 0 || device >= ndev) return qvti_fail(QVTI_E_ARG, "device %d of %d", device, ndev);
    HIPCHK(hipSetDevice(device));

    qvti_ctx* c = new qvti_ctx();
    c->device = device;
    c->g = *g;
    c->S = max_shots;
    c->N = (long long)g->nx * g->ny * g->nz;
    c->Nb = blk_len(c->N);
    {
        // split variants: the stored choice for this grid (qvti_tune_split,
        // qvti_load_wisdom), else the default
        const std::array<int, 3> sp = stored_split(*g);
        for (int a = 0; a < 3; ++a) {
            c->split[a] = len_ops(dims[a], sp[a]) ? sp[a] : 0;
            c->ops[a] = len_ops(dims[a], c->split[a]);
        }
    }
    for (int a = 0; a < 3; ++a) {
        c->snap_lo[a] = g->pml;
        c->snap_hi[a] = dims[a] - g->pml;
    }
    c->nrows = (long long)g->nx * g->ny;
    // When a z row does not end on a 128-byte line, neighbouring row tiles
    // and neighbouring planes share lines (and 64-cell blocks): run 8 row
    // tiles and 2 planes per XCD so those are fetched once.  Measured at
    // 400x400x200 (800-byte rows): str_z 2472 -> 2380 us, str_x 1066 ->
    // 1036 us; at 256^3 (aligned rows) one tile and one plane per XCD is best.
    if ((g->nz * 4) % 128 != 0) {
        c->run_rows = 8;
        c->run_planes = 2;
    }
    int rc = QVTI_OK;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return qvti_fail(QVTI_E_DEVICE, "hipStreamCreate failed");
    }
    const float ds[3] = {g->dx, g->dy, g->dz};
    for (int a = 0; a < 3 && !rc; ++a) {
        AxisDev& A = c->ax[a];
        A.n = dims[a];
        A.d = ds[a];
        std::vector<float2> tw(A.n), dp, dm;
        for (int t = 0; t < A.n; ++t) {
            const double ang = -2.0 * M_PI * t / A.n;
            tw[t] = make_float2((float)cos(ang), (float)sin(ang));
        }
        std::vector<float> kf, pa, pb, pah, pbh;
        axis_symbols(A.n, A.d, dp, dm, kf);
        pml_profile(A.n, g->pml, A.d, g->f0, pa, pb, pah, pbh);
        rc = rc ? rc : upload(c, &A.tw, tw);
        rc = rc ? rc : upload(c, &A.dp, dp);
        rc = rc ? rc : upload(c, &A.dm, dm);
        rc = rc ? rc : upload(c, &A.kf, kf);
        rc = rc ? rc : upload(c, &A.a, pa);
        rc = rc ? rc : upload(c, &A.b, pb);
        rc = rc ? rc : upload(c, &A.ah, pah);
        rc = rc ? rc : upload(c, &A.bh, pbh);
    }
    const long long N = c->N, S = c->S;
    const int p2 = 2 * g->pml;
    c->slab[0] = (long long)p2 * g->ny * g->nz;
    c->slab[1] = (long long)g->nx * p2 * g->nz;
    c->slab[2] = (long long)g->nx * g->ny * p2;
    const long long Nb = c->Nb;
    rc = rc ? rc : dalloc(c, (void**)&c->coef, sizeof(float) * (kG1 + (g->vti ? kG2 : 0)) * Nb);
    rc = rc ? rc : dalloc(c, (void**)&c->buoy, sizeof(float) * 3 * N);
    rc = rc ? rc : dalloc(c, (void**)&c->v, sizeof(float) * 3 * S * N);
    rc = rc ? rc : dalloc(c, (void**)&c->sig, sizeof(float) * 6 * S * N);
    rc = rc ? rc : dalloc(c, (void**)&c->dv, sizeof(float) * 3 * S * N);
    rc = rc ? rc : dalloc(c, (void**)&c->H, sizeof(float) * 9 * S * N);
    rc = rc ? rc : dalloc(c, (void**)&c->bufA, sizeof(float) * 15 * S * Nb);
    rc = rc ? rc : dalloc(c, (void**)&c->bufB, sizeof(float) * 15 * S * Nb);
    for (int a = 0; a < 3 && !rc; ++a) {
        rc = dalloc(c, (void**)&c->phiv[a], sizeof(float) * 3 * S * c->slab[a]);
        rc = rc ? rc : dalloc(c, (void**)&c->phis[a], sizeof(float) * c->nphis[a] * S * c->slab[a]);
    }
    rc = rc ? rc : dalloc(c, (void**)&c->d_it, sizeof(int));
    rc = rc ? rc : dalloc(c, (void**)&c->d_probe, sizeof(int));
    rc = rc ? rc : dalloc(c, (void**)&c->d_src_kind, sizeof(int) * S);
    rc = rc ? rc : dalloc(c, (void**)&c->d_src_ip, sizeof(long long) * S);
    rc = rc ? rc : dalloc(c, (void**)&c->d_rec_off, sizeof(int) * (S + 1));
    if (!rc) rc = build_launch_config(c);
    if (rc) {
        std::string keep = qvti_last_error();
        release(c);
        delete c;
        qvti_fail(rc, "%s", keep.c_str());
        return rc;
    }
    *out = c;
    return QVTI_OK;
}

extern "C" int qvti_destroy(qvti_ctx* c)
{
    if (!c) return QVTI_OK;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    release(c);
    delete c;
    return QVTI_OK;
}


static int build_launch_config(qvti_ctx* c)
{
    const qvti_grid& g = c->g;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long long sx = (long long)ny * nz, sy = nz, N = c->N;
    const AxisDev &X = c->ax[0], &Y = c->ax[1], &Z = c->ax[2];
    for (int a = 0; a < 2; ++a) c->P[a] = c->ops[a]->P[0];
    c->P[2] = c->ops[2]->Prow;
    const int Px = c->P[0], Py = c->P[1], Pz = c->P[2];

    float* acc[3] = {fld(c->bufA, 0, c), fld(c->bufA, 1, c), fld(c->bufA, 2, c)};
    float* vv[3] = {fld(c->v, 0, c), fld(c->v, 1, c), fld(c->v, 2, c)};
    float* dvv[3] = {fld(c->dv, 0, c), fld(c->dv, 1, c), fld(c->dv, 2, c)};
    // sig order: pxx pyy pzz pxy pxz pyz
    float* sg[6];
    for (int k = 0; k < 6; ++k) sg[k] = fld(c->sig, k, c);
    float *pxx = sg[0], *pyy = sg[1], *pzz = sg[2], *pxy = sg[3], *pxz = sg[4], *pyz = sg[5];

    auto strided_common = [&](StridedArgs& a, int axis, int P) {
        const AxisDev& A = c->ax[axis];
        a.nx = nx; a.ny = ny; a.nz = nz; a.pml = g.pml;
        a.sx = sx; a.sy = sy;
        a.axis = axis;
        a.ztiles = (nz + 2 * P - 1) / (2 * P);
        a.nplanes = axis == 0 ? ny : nx;
        a.run = 1;
        a.dt = g.dt;
        a.shot_stride = N;
        a.slab_stride = c->slab[axis];
        a.tw = A.tw;
        a.sym[0] = A.dp;
        a.sym[1] = A.dm;
        a.pa = A.a; a.pb = A.b; a.pah = A.ah; a.pbh = A.bh;
    };
    auto job = [](const float* in, float* out, const float* accp, float* phi, int sym, int half) {
        DJob j;
        j.in = in; j.out = out; j.acc = accp; j.phi = phi; j.sym = sym; j.half = half;
        return j;
    };
    // ---- velocity x: pxx (D-, half) -> acc_x, pxy (D+) -> acc_y, pxz (D+) -> acc_z
    StridedArgs& vx = c->a_vx;
    strided_common(vx, 0, Px);
    vx.mode = MODE_VEL_FIRST;
    vx.njobs = 3;
    vx.job[0] = job(pxx, acc[0], nullptr, nullptr, 1, 1);
    vx.job[1] = job(pxy, acc[1], nullptr, nullptr, 0, 0);
    vx.job[2] = job(pxz, acc[2], nullptr, nullptr, 0, 0);
    vx.tick = c->d_it;  // the first kernel of a step advances the step counter
    for (int k = 0; k < 3; ++k) vx.job[k].phi = c->phiv[0] + (long long)k * c->S * c->slab[0];  // [field][shot][slab]

    // ---- velocity y: pxy (D+) -> acc_x, pyy (D-, half) -> acc_y, pyz (D+) -> acc_z
    StridedArgs& vy = c->a_vy;
    strided_common(vy, 1, Py);
    vy.mode = MODE_VEL_ACC;
    vy.njobs = 3;
    vy.job[0] = job(pxy, acc[0], acc[0], nullptr, 0, 0);
    vy.job[1] = job(pyy, acc[1], acc[1], nullptr, 1, 1);
    vy.job[2] = job(pyz, acc[2], acc[2], nullptr, 0, 0);
    for (int k = 0; k < 3; ++k) vy.job[k].phi = c->phiv[1] + (long long)k * c->S * c->slab[1];

    // ---- 3-D transform buffers: P_c = FFT(v_c + i dv_c) (bufB as complex),
    // Q1_c, Q3_c (bufA as complex)
    float2* P[3];
    float2* Q1[3];
    float2* Q3[3];
    for (int k = 0; k < 3; ++k) {
        P[k] = reinterpret_cast<float2*>(fld(c->bufB, 2 * k, c));
        Q1[k] = reinterpret_cast<float2*>(fld(c->bufA, 2 * k, c));
        Q3[k] = reinterpret_cast<float2*>(fld(c->bufA, 6 + 2 * k, c));
    }
    // ---- velocity z + update + forward z-FFT: pxz (D+) -> vx, pyz (D+) -> vy, pzz (D-, half) -> vz
    VelZArgs& vz = c->a_vz;
    vz.nx = nx; vz.ny = ny; vz.nz = nz; vz.pml = g.pml;
    vz.nrows = c->nrows;
    vz.dt = g.dt;
    vz.shot_stride = N;
    vz.slab_stride = c->slab[2];
    vz.tw = Z.tw; vz.symp = Z.dp; vz.symm = Z.dm;
    vz.pa = Z.a; vz.pb = Z.b; vz.pah = Z.ah; vz.pbh = Z.bh;
    vz.pxz = pxz; vz.pyz = pyz; vz.pzz = pzz;
    for (int k = 0; k < 3; ++k) {
        vz.acc[k] = acc[k];
        vz.v[k] = vv[k];
        vz.dv[k] = dvv[k];
        vz.P[k] = P[k];
        vz.buoy[k] = c->buoy + (long long)k * N;
        vz.phi[k] = c->phiv[2] + (long long)k * c->S * c->slab[2];
    }
    vz.src_kind = c->d_src_kind;
    vz.src_ip = c->d_src_ip;
    vz.rik = nullptr;  // set at begin
    vz.it = c->d_it;
    c->g_vz = dim3((unsigned)((c->nrows + Pz - 1) / Pz), 1, 1);
    vz.ntiles = (int)c->g_vz.x;
    vz.nshots = 1;
    c->g_zr = c->g_vz;

    // H: [A1 v x,y,z][A1 dv x,y,z][A3 v x,y,z]
    float* H1v[3];
    float* H1dv[3];
    float* H3v[3];
    for (int k = 0; k < 3; ++k) {
        H1v[k] = fld(c->H, k, c);
        H1dv[k] = fld(c->H, 3 + k, c);
        H3v[k] = fld(c->H, 6 + k, c);
    }
    CStridedArgs& yf = c->a_yf;
    yf.nz = nz; yf.sx = sx; yf.sy = sy;
    yf.axis = 1;
    yf.ztiles = (nz + Py - 1) / Py;
    yf.nplanes = nx;
    yf.run = 1;
    yf.shot_stride = N;
    yf.tw = Y.tw;
    yf.nfields = 3;
    for (int k = 0; k < 3; ++k) yf.f[k] = P[k];

    XPassArgs& xp = c->a_xp;
    xp.ny = ny; xp.nz = nz; xp.sx = sx; xp.sy = sy;
    xp.ztiles = (nz + Px - 1) / Px;
    xp.run = 1;
    xp.shot_stride = N;
    xp.tw = X.tw;
    xp.kfx = X.kf; xp.kfy = Y.kf; xp.kfz = Z.kf;
    xp.invN = 1.0 / (double)N;
    xp.ncomp = 3;
    // P0 = F(vx + i vy), P1 = F(dvx + i dvy), P2 = F(vz + i dvz): A1 on all
    // three, |k| on P0 and P2 (A3 vx + i A3 vy, A3 vz + i unused)
    for (int k = 0; k < 3; ++k) { xp.P[k] = P[k]; xp.Q1[k] = Q1[k]; }
    xp.Q3[0] = Q3[0];
    xp.Q3[1] = nullptr;
    xp.Q3[2] = Q3[1];
    // k-space mode: A2 F P_c -> Q2_c in bufB fields 6-11 (free once the
    // x-pass has read P; the y-derivatives overwrite them after z_inv)
    float2* Q2[3] = {nullptr, nullptr, nullptr};
    float* H2v[3] = {nullptr, nullptr, nullptr};
    float* H2dv[3] = {nullptr, nullptr, nullptr};
    xp.ks_vh = xp.ks_k0 = xp.ks_kc = xp.ks_pw = 0.f;
    if (c->ks) {
        for (int k = 0; k < 3; ++k) {
            Q2[k] = reinterpret_cast<float2*>(fld(c->bufB, 6 + 2 * k, c));
            H2v[k] = fld(c->H2, k, c);
            H2dv[k] = fld(c->H2, 3 + k, c);
            xp.Q2[k] = Q2[k];
        }
        xp.ks_vh = (float)(0.5 * (double)c->ks_vref * (double)g.dt);
        if (c->ks_fcut > 0.f) {
            const double kc = 2.0 * kRefPi * (double)c->ks_fcut / (double)c->ks_vmax;
            const double k0 = (1.0 - (double)c->ks_taper) * kc;
            xp.ks_kc = (float)kc;
            xp.ks_k0 = (float)k0;
            xp.ks_pw = kc > k0 ? (float)(M_PI / (kc - k0)) : 0.f;
        }
    }

    CStridedArgs& yi = c->a_yi;
    yi = yf;
    yi.nfields = 5;
    for (int k = 0; k < 3; ++k) yi.f[k] = Q1[k];
    yi.f[3] = Q3[0];
    yi.f[4] = Q3[1];
    if (c->ks) {
        yi.nfields = 8;
        for (int k = 0; k < 3; ++k) yi.f[5 + k] = Q2[k];
    }

    ZRowsArgs& zi = c->a_zi;
    zi.nrows = c->nrows;
    zi.shot_stride = N;
    zi.tw = Z.tw;
    zi.mode = 1;
    zi.dir = 1;
    zi.nfields = 5;
    zi.cin[0] = Q1[0]; zi.re_out[0] = H1v[0]; zi.im_out[0] = H1v[1];    // A1 vx, A1 vy
    zi.cin[1] = Q1[1]; zi.re_out[1] = H1dv[0]; zi.im_out[1] = H1dv[1];  // A1 dvx, A1 dvy
    zi.cin[2] = Q1[2]; zi.re_out[2] = H1v[2]; zi.im_out[2] = H1dv[2];   // A1 vz, A1 dvz
    zi.cin[3] = Q3[0]; zi.re_out[3] = H3v[0]; zi.im_out[3] = H3v[1];    // A3 vx, A3 vy
    zi.cin[4] = Q3[1]; zi.re_out[4] = H3v[2]; zi.im_out[4] = nullptr;   // A3 vz
    if (c->ks) {
        zi.nfields = 8;
        zi.cin[5] = Q2[0]; zi.re_out[5] = H2v[0]; zi.im_out[5] = H2v[1];    // A2 vx, A2 vy
        zi.cin[6] = Q2[1]; zi.re_out[6] = H2dv[0]; zi.im_out[6] = H2dv[1];  // A2 dvx, A2 dvy
        zi.cin[7] = Q2[2]; zi.re_out[7] = H2v[2]; zi.im_out[7] = H2dv[2];   // A2 vz, A2 dvz
    }

    // ---- stress x / y derivatives of the 15 family fields
    const float* fam[3][5];
    for (int k = 0; k < 3; ++k) {
        fam[k][0] = H1v[k]; fam[k][1] = vv[k]; fam[k][2] = H3v[k]; fam[k][3] = H1dv[k]; fam[k][4] = dvv[k];
        if (c->ks) { fam[k][1] = H2v[k]; fam[k][4] = H2dv[k]; }
    }
    // x/y derivatives of the stress step: blocked groups of 15 fields per
    // shot in bufA (x) and bufB (y); job k writes field k of the group
    float* DX[15];
    float* DY[15];
    // x-derivatives: three blocked groups of 5 (ex, gx, hx), each [shot][N/64][5][64];
    // y-derivatives: one blocked group of 15
    for (int k = 0; k < 15; ++k) {
        DX[k] = c->bufA + (long long)(k / 5) * 5 * c->S * c->Nb + (long long)(k % 5) * kBlk;
        DY[k] = c->bufB + (long long)k * kBlk;
    }
    StridedArgs& sxa = c->a_sx;
    strided_common(sxa, 0, Px);
    sxa.mode = MODE_STORE;
    sxa.njobs = 15;
    sxa.bnf = 5;
    sxa.blk_stride = c->Nb;
    for (int j = 0; j < 5; ++j) {
        sxa.job[j] = job(fam[0][j], DX[j], nullptr, nullptr, 0, 0);          // ex: dx vx-family, D+
        sxa.job[5 + j] = job(fam[2][j], DX[5 + j], nullptr, nullptr, 1, 0);   // gx: dx vz-family, D-
        sxa.job[10 + j] = job(fam[1][j], DX[10 + j], nullptr, nullptr, 1, 0); // hx: dx vy-family, D-
    }
    StridedArgs& sya = c->a_sy;
    strided_common(sya, 1, Py);
    sya.mode = MODE_STORE;
    sya.njobs = 15;
    sya.bnf = kG1;
    sya.blk_stride = c->Nb;
    for (int j = 0; j < 5; ++j) {
        sya.job[j] = job(fam[1][j], DY[j], nullptr, nullptr, 0, 0);          // ey: dy vy-family, D+
        sya.job[5 + j] = job(fam[2][j], DY[5 + j], nullptr, nullptr, 1, 0);   // gy: dy vz-family, D-
        sya.job[10 + j] = job(fam[0][j], DY[10 + j], nullptr, nullptr, 1, 0); // hy: dy vx-family, D-
    }

    // ---- stress z + update
    StressZArgs& sz = c->a_sz;
    sz.nx = nx; sz.ny = ny; sz.nz = nz; sz.pml = g.pml;
    sz.nrows = c->nrows;
    sz.dt = g.dt;
    sz.shot_stride = N;
    sz.slabx_stride = c->slab[0]; sz.slaby_stride = c->slab[1]; sz.slabz_stride = c->slab[2];
    sz.tw = Z.tw; sz.symp = Z.dp; sz.symm = Z.dm;
    sz.pax = X.a; sz.pbx = X.b; sz.paxh = X.ah; sz.pbxh = X.bh;
    sz.pay = Y.a; sz.pby = Y.b; sz.payh = Y.ah; sz.pbyh = Y.bh;
    sz.paz = Z.a; sz.pbz = Z.b; sz.pazh = Z.ah; sz.pbzh = Z.bh;
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 5; ++j) sz.fam[k][j] = fam[k][j];
    sz.dxb = c->bufA;
    sz.dx_group = 5LL * c->S * c->Nb;
    sz.blk_stride = c->Nb;
    sz.dyb = c->bufB;
    sz.g1 = c->coef;
    sz.g2 = g.vti ? c->coef + (long long)kG1 * c->Nb : nullptr;
    for (int k = 0; k < 6; ++k) sz.sig[k] = sg[k];
    for (int k = 0; k < 3; ++k) sz.vel[k] = vv[k];
    for (int k = 0; k < 5; ++k) {
        sz.phx[k] = c->phis[0] + (long long)k * c->S * c->slab[0];
        sz.phy[k] = c->phis[1] + (long long)k * c->S * c->slab[1];
    }
    for (int k = 0; k < 4; ++k) sz.phz[k] = c->phis[2] + (long long)k * c->S * c->slab[2];
    sz.src_kind = c->d_src_kind;
    sz.src_ip = c->d_src_ip;
    sz.rik = nullptr;
    sz.it = c->d_it;
    const int R = c->ops[2]->Rsz;
    c->g_sz = dim3((unsigned)((c->nrows + R - 1) / R), 1, 1);
    sz.ntiles = (int)c->g_sz.x;
    sz.nshots = 1;

    RecArgs& rec = c->a_rec;
    rec.vx = vv[0]; rec.vy = vv[1]; rec.vz = vv[2];
    rec.shot_stride = N;
    rec.rec_off = c->d_rec_off;
    rec.it = c->d_it;
    (void)pyy;
    for (int k = 0; k < KC_COUNT; ++k) c->pw[k] = c->jz[k] = 1;
    apply_tiles(c);
    return QVTI_OK;
}

// ------------------------------------------------------------ model ------
extern "C" int qvti_set_model(qvti_ctx* c, const float* velp, const float* gama_p, const float* vels,
                              const float* gama_s, const float* rho, const float* eps, const float* delta)
{
    if (!c) return qvti_fail(QVTI_E_ARG, "NULL context");
    if (!velp || !gama_p || !vels || !gama_s || !rho) return qvti_fail(QVTI_E_ARG, "model array is NULL");
    if (c->g.vti && (!eps || !delta)) return qvti_fail(QVTI_E_ARG, "VTI grid needs eps and delta");
    HIPCHK(hipSetDevice(c->device));
    const long long N = c->N;
    const size_t B = sizeof(float) * N;
    // The model is staged in the spectral scratch buffer (15 fields per shot,
    // free between runs and ordered on the context's stream), so setting a
    // model allocates nothing and never waits for other contexts' work.
    float* tmp = c->bufA;
    const int nin = c->g.vti ? 7 : 5;
    const float* srcs[7] = {velp, gama_p, vels, gama_s, rho, eps, delta};
    int rc = QVTI_OK;
    for (int k = 0; k < nin && !rc; ++k)
        if (hipMemcpyAsync(tmp + k * N, srcs[k], B, hipMemcpyHostToDevice, c->stream) != hipSuccess)
            rc = qvti_fail(QVTI_E_DEVICE, "model upload failed");
    if (!rc) {
        CoefArgs ca;
        ca.N = N;
        ca.vti = c->g.vti;
        ca.media = c->media;
        ca.w0 = (double)c->g.w0;
        ca.dt = (double)c->g.dt;
        ca.vp = tmp; ca.gp = tmp + N; ca.vs = tmp + 2 * N; ca.gs = tmp + 3 * N; ca.rho = tmp + 4 * N;
        ca.eps = c->g.vti ? tmp + 5 * N : nullptr;
        ca.del = c->g.vti ? tmp + 6 * N : nullptr;
        ca.g1 = c->coef;
        ca.g2 = c->g.vti ? c->coef + (long long)kG1 * c->Nb : nullptr;
        const unsigned nb = (unsigned)std::min<long long>((N + kThreads - 1) / kThreads, 65535LL * 4);
        hipLaunchKernelGGL(k_coefficients, dim3(nb), dim3(kThreads), 0, c->stream, ca);
        AvgArgs aa;
        aa.nx = c->g.nx; aa.ny = c->g.ny; aa.nz = c->g.nz;
        aa.sx = (long long)c->g.ny * c->g.nz; aa.sy = c->g.nz; aa.N = N;
        aa.dt = c->g.dt;
        aa.g1 = c->coef;
        aa.rho = tmp + 4 * N;
        aa.buoy = c->buoy;
        hipLaunchKernelGGL(k_averages, dim3(nb), dim3(kThreads), 0, c->stream, aa);
        if (hipStreamSynchronize(c->stream) != hipSuccess || hipGetLastError() != hipSuccess)
            rc = qvti_fail(QVTI_E_DEVICE, "model setup kernels failed");
    }
    if (!rc) c->has_model = true;
    return rc;
}

// ------------------------------------------------------------ running ----
static int record_event(qvti_ctx* c, int cls)
{
    if (!c->timing || c->ev_used >= c->ev_max) return -1;
    const int k = c->ev_used++;
    hipEventRecord(c->ev[k], c->stream);
    c->ev_class[k] = cls;
    return k;
}

// Workgroups per row tile of the final inverse z pass: 1 puts all fields in
// one workgroup, any other value one field per workgroup.
#ifndef QVTI_ZINV_SPLIT
#define QVTI_ZINV_SPLIT 5
#endif

// 1-D grid of the XCD-aware row kernels (tile_shot): ceil(ntiles/(8 run))*8 run*S.
static unsigned xcd_grid(int ntiles, unsigned S, int run)
{
    const int g = 8 * run;
    return (unsigned)((ntiles + g - 1) / g) * (unsigned)g * S;
}

// Launch one kernel of class cls on the context's stream.
static void launch_class(qvti_ctx* c, int cls)
{
    hipStream_t st = c->stream;
    const unsigned S = (unsigned)c->nshots;
    const LenOps &X = *c->ops[0], &Y = *c->ops[1], &Z = *c->ops[2];
    auto strided = [&](const dim3& g) { return dim3(g.x, S, (unsigned)c->jz[cls]); };
    switch (cls) {
    case KC_VEL_X:
        X.deriv[MODE_VEL_FIRST][pw_index(X, c->pw[cls])](strided(c->g_vx), st, c->a_vx);
        break;
    case KC_VEL_Y:
        Y.deriv[MODE_VEL_ACC][pw_index(Y, c->pw[cls])](strided(c->g_vy), st, c->a_vy);
        break;
    case KC_VEL_Z:
        Z.vel_z[c->a_vz.inj[0] ? 1 : 0](dim3(xcd_grid(c->a_vz.ntiles, S, c->a_vz.run)), st, c->a_vz);
        break;
    case KC_Y_FWD:
        Y.cfft[0][pw_index(Y, c->pw[cls])](strided(c->g_cy), st, c->a_yf);
        break;
    case KC_XPASS:
        X.xpass[c->ks ? 1 : 0][pw_index(X, c->pw[cls])](strided(c->g_xp), st, c->a_xp);
        break;
    case KC_Y_INV:
        Y.cfft[1][pw_index(Y, c->pw[cls])](strided(c->g_yi), st, c->a_yi);
        break;
    case KC_Z_INV: {
        const unsigned zs = QVTI_ZINV_SPLIT == 1 ? 1u : (unsigned)c->a_zi.nfields;
        Z.zrows(dim3(c->g_zr.x, S, zs), st, c->a_zi);
        break;
    }
    case KC_STR_X:
        X.deriv[MODE_STORE][pw_index(X, c->pw[cls])](strided(c->g_sx), st, c->a_sx);
        break;
    case KC_STR_Y:
        Y.deriv[MODE_STORE][pw_index(Y, c->pw[cls])](strided(c->g_sy), st, c->a_sy);
        break;
    case KC_STR_Z:
        Z.stress_z[c->g.vti ? 1 : 0](dim3(xcd_grid(c->a_sz.ntiles, S, c->a_sz.run)), st, c->a_sz);
        break;
    case KC_RECORD:
        hipLaunchKernelGGL(k_record, dim3(c->g_rec.x, S), dim3(kThreads), 0, st, c->a_rec);
        break;
    default:
        break;
    }
}

// Enqueue one time step (11 launches; vel_x also advances the step counter).
// With timing on, each launch is bracketed by an event pair.
static void enqueue_step(qvti_ctx* c)
{
    static const int order[] = {KC_VEL_X, KC_VEL_Y, KC_VEL_Z, KC_Y_FWD, KC_XPASS,
                                KC_Y_INV, KC_Z_INV, KC_STR_X, KC_STR_Y, KC_STR_Z, KC_RECORD};
    for (int cls : order) {
        if (cls == KC_RECORD && c->total_rec == 0) continue;
        const int e0 = c->timing ? record_event(c, cls) : -1;
        launch_class(c, cls);
        if (e0 >= 0) record_event(c, -1 - cls);
    }
}

// Geometry of the strided-axis launches for the current tile widths.
static void apply_tiles(qvti_ctx* c)
{
    const int nx = c->g.nx, ny = c->g.ny, nz = c->g.nz;
    auto zt = [&](int axis, int cls, int cols) {
        const int w = tile_w(c->ops[axis], c->pw[cls]).P * cols;  // z-columns per tile
        return (nz + w - 1) / w;
    };
    // XCD-aware grid of nplanes planes of ztiles tiles, `run` consecutive
    // planes per XCD (plane_tile())
    auto grid = [&](int nplanes, int ztiles, int run) {
        const int g = 8 * run;
        return dim3((unsigned)((nplanes + g - 1) / g) * (unsigned)g * ztiles);
    };
    const int rp = c->run_planes;
    c->a_vx.run = c->a_sx.run = c->a_vy.run = c->a_sy.run = c->a_yf.run = c->a_yi.run = c->a_xp.run = rp;
    c->a_vx.ztiles = zt(0, KC_VEL_X, 2); c->g_vx = grid(ny, c->a_vx.ztiles, rp);
    c->a_sx.ztiles = zt(0, KC_STR_X, 2); c->g_sx = grid(ny, c->a_sx.ztiles, rp);
    c->a_vy.ztiles = zt(1, KC_VEL_Y, 2); c->g_vy = grid(nx, c->a_vy.ztiles, rp);
    c->a_sy.ztiles = zt(1, KC_STR_Y, 2); c->g_sy = grid(nx, c->a_sy.ztiles, rp);
    c->a_yf.ztiles = zt(1, KC_Y_FWD, 1); c->g_cy = grid(nx, c->a_yf.ztiles, rp);
    c->a_yi.ztiles = zt(1, KC_Y_INV, 1); c->g_yi = grid(nx, c->a_yi.ztiles, rp);
    c->a_xp.ztiles = zt(0, KC_XPASS, 1); c->g_xp = grid(ny, c->a_xp.ztiles, rp);
    c->a_vz.run = c->a_sz.run = c->run_rows;
}

// Fields (jobs) one launch of a strided kernel class processes.
static int class_fields(const qvti_ctx* c, int cls)
{
    switch (cls) {
    case KC_VEL_X: case KC_VEL_Y: case KC_Y_FWD: case KC_XPASS: return 3;
    case KC_Y_INV: return c->a_yi.nfields;  // 5, or 8 in k-space mode
    case KC_STR_X: case KC_STR_Y: return 15;
    default: return 1;
    }
}

// Autotuner: time every available tile width of each strided-axis kernel
// class on this context's grid and shot count, keep the fastest.  Runs on the
// context's own buffers before a run starts: the velocity-step kernels see
// zero stresses (so the C-PML memory variables stay zero) and every other
// tuned kernel only writes scratch buffers that a step overwrites first.
static int autotune_tiles(qvti_ctx* c)
{
    static const int classes[] = {KC_VEL_X, KC_VEL_Y, KC_Y_FWD, KC_XPASS, KC_Y_INV, KC_STR_X, KC_STR_Y};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        hipEventDestroy(e0);
        return qvti_fail(QVTI_E_DEVICE, "hipEventCreate failed");
    }
    // the timing launches must not advance the step counter
    int* const tick = c->a_vx.tick;
    c->a_vx.tick = nullptr;
    int rc = QVTI_OK;
    for (int cls : classes) {
        const LenOps* o = (cls == KC_VEL_X || cls == KC_XPASS || cls == KC_STR_X) ? c->ops[0] : c->ops[1];
        const int nf = class_fields(c, cls);
        int best_pw = 1, best_jz = 1;
        float best = 1e30f;
        for (int pw : {1, 2, 4}) {
            if (tile_w(o, pw).pw != pw) continue;  // width not available for this length
            for (int jz : {1, nf}) {
                c->pw[cls] = pw;
                c->jz[cls] = jz;
                apply_tiles(c);
                launch_class(c, cls);  // warm-up
                hipEventRecord(e0, c->stream);
                for (int r = 0; r < 3; ++r) launch_class(c, cls);
                hipEventRecord(e1, c->stream);
                float ms = 0.f;
                if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                    rc = qvti_fail(QVTI_E_DEVICE, "autotune timing failed");
                    break;
                }
                if (ms < best) {
                    best = ms;
                    best_pw = pw;
                    best_jz = jz;
                }
            }
            if (rc) break;
        }
        c->pw[cls] = best_pw;
        c->jz[cls] = best_jz;
        if (rc) break;
    }
    c->a_vx.tick = tick;
    apply_tiles(c);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (!rc && hipGetLastError() != hipSuccess) rc = qvti_fail(QVTI_E_DEVICE, "autotune launch failed");
    return rc;
}

// Cells of the snapshot box.
static long long snap_cells(const qvti_ctx* c)
{
    return (long long)(c->snap_hi[0] - c->snap_lo[0]) * (c->snap_hi[1] - c->snap_lo[1]) *
           (c->snap_hi[2] - c->snap_lo[2]);
}

static const float* snap_source(qvti_ctx* c)
{
    const int f = c->snap_field;
    if (f < 3) return fld(c->v, f, c);
    return fld(c->sig, f - 3, c);
}

// One step plus the snapshot copy of shot 0 when snapshots are enabled.
static void enqueue_step_with_snap(qvti_ctx* c)
{
    enqueue_step(c);
    if (c->snap_every > 0 && c->d_snap) {
        const int e0 = c->timing ? record_event(c, KC_SNAP) : -1;
        SnapArgs sa;
        sa.f = snap_source(c);
        sa.ny = c->g.ny; sa.nz = c->g.nz;
        for (int a = 0; a < 3; ++a) { sa.lo[a] = c->snap_lo[a]; sa.hi[a] = c->snap_hi[a]; }
        sa.every = c->snap_every;
        sa.it = c->d_it;
        sa.out = c->d_snap;
        sa.cap = c->snap_lim;
        hipLaunchKernelGGL(k_snapshot, dim3(256), dim3(kThreads), 0, c->stream, sa);
        if (e0 >= 0) record_event(c, -1 - KC_SNAP);
    }
}

static int capture_graph(qvti_ctx* c)
{
    if (c->gexec) {
        hipGraphExecDestroy(c->gexec);
        c->gexec = nullptr;
    }
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const bool t = c->timing;
    c->timing = false;
    enqueue_step_with_snap(c);
    c->timing = t;
    hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (e != hipSuccess) return qvti_fail(QVTI_E_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&c->gexec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) {
        c->gexec = nullptr;
        return qvti_fail(QVTI_E_DEVICE, "graph instantiate failed: %s", hipGetErrorString(e));
    }
    return QVTI_OK;
}

extern "C" int qvti_begin(qvti_ctx* c, int nt, const float* rik, const qvti_shot* shots, int nshots)
{
    if (!c || !rik || !shots) return qvti_fail(QVTI_E_ARG, "NULL argument");
    if (!c->has_model) return qvti_fail(QVTI_E_STATE, "qvti_set_model must be called first");
    if (nt < 1) return qvti_fail(QVTI_E_ARG, "nt must be >= 1");
    if (nshots < 1 || nshots > c->S) return qvti_fail(QVTI_E_ARG, "nshots %d not in [1, %d]", nshots, c->S);
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    const qvti_grid& g = c->g;
    std::vector<int> kinds(nshots), off(nshots + 1, 0);
    std::vector<long long> sip(nshots), rip;
    for (int s = 0; s < nshots; ++s) {
        const qvti_shot& sh = shots[s];
        if (sh.kind < QVTI_SRC_FX || sh.kind > QVTI_SRC_EXPLOSIVE)
            return qvti_fail(QVTI_E_ARG, "shot %d: unknown source kind %d", s, sh.kind);
        if (sh.sx < 0 || sh.sx >= g.nx || sh.sy < 0 || sh.sy >= g.ny || sh.sz < 0 || sh.sz >= g.nz)
            return qvti_fail(QVTI_E_ARG, "shot %d: source cell (%d,%d,%d) outside the grid", s, sh.sx, sh.sy, sh.sz);
        if (sh.nrec < 0 || (sh.nrec > 0 && (!sh.rx || !sh.ry || !sh.rz)))
            return qvti_fail(QVTI_E_ARG, "shot %d: receiver arrays", s);
        kinds[s] = sh.kind;
        sip[s] = ((long long)sh.sx * g.ny + sh.sy) * g.nz + sh.sz;
        for (int r = 0; r < sh.nrec; ++r) {
            if (sh.rx[r] < 0 || sh.rx[r] >= g.nx || sh.ry[r] < 0 || sh.ry[r] >= g.ny || sh.rz[r] < 0 || sh.rz[r] >= g.nz)
                return qvti_fail(QVTI_E_ARG, "shot %d receiver %d outside the grid", s, r);
            rip.push_back(((long long)sh.rx[r] * g.ny + sh.ry[r]) * g.nz + sh.rz[r]);
        }
        off[s + 1] = off[s] + sh.nrec;
    }
    // From here on the previous run's buffers may be freed: the context is
    // not steppable again until this call succeeds (its captured graph holds
    // the previous run's kernel arguments).
    c->begun = 0;
    c->steps_done = 0;
    if (c->gexec) {
        hipGraphExecDestroy(c->gexec);
        c->gexec = nullptr;
    }
    // (re)allocate run buffers
    auto realloc_dev = [&](void** p, size_t bytes) -> int {
        if (*p) {
            auto it = std::find(c->allocs.begin(), c->allocs.end(), *p);
            if (it != c->allocs.end()) c->allocs.erase(it);
            hipFree(*p);
            *p = nullptr;
        }
        if (bytes == 0) return QVTI_OK;
        return dalloc(c, p, bytes);
    };
    int rc = QVTI_OK;
    if (c->nt != nt || !c->d_rik) rc = realloc_dev((void**)&c->d_rik, sizeof(float) * nt);
    const long long rec_need = 3LL * off[nshots] * nt;
    if (!rc && rec_need > c->rec_cap) {
        rc = realloc_dev((void**)&c->d_rec, sizeof(float) * rec_need);
        c->rec_cap = rc ? 0 : rec_need;
    }
    if (!rc) rc = realloc_dev((void**)&c->d_rec_ip, sizeof(long long) * std::max<size_t>(rip.size(), 1));
    if (rc) return rc;
    c->nt = nt;
    c->nshots = nshots;
    c->total_rec = off[nshots];
    c->steps_done = 0;
    HIPCHK(hipMemcpyAsync(c->d_rik, rik, sizeof(float) * nt, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_src_kind, kinds.data(), sizeof(int) * nshots, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_src_ip, sip.data(), sizeof(long long) * nshots, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_rec_off, off.data(), sizeof(int) * (nshots + 1), hipMemcpyHostToDevice, c->stream));
    if (!rip.empty())
        HIPCHK(hipMemcpyAsync(c->d_rec_ip, rip.data(), sizeof(long long) * rip.size(), hipMemcpyHostToDevice, c->stream));
    // zero the state that is accumulated rather than overwritten
    const long long N = c->N, S = c->S;
    HIPCHK(hipMemsetAsync(c->v, 0, sizeof(float) * 3 * S * N, c->stream));
    HIPCHK(hipMemsetAsync(c->sig, 0, sizeof(float) * 6 * S * N, c->stream));
    for (int a = 0; a < 3; ++a) {
        HIPCHK(hipMemsetAsync(c->phiv[a], 0, sizeof(float) * 3 * S * c->slab[a], c->stream));
        HIPCHK(hipMemsetAsync(c->phis[a], 0, sizeof(float) * c->nphis[a] * S * c->slab[a], c->stream));
    }
    if (c->d_rec) HIPCHK(hipMemsetAsync(c->d_rec, 0, sizeof(float) * rec_need, c->stream));
    const int minus1 = -1, zero = 0;
    HIPCHK(hipMemcpyAsync(c->d_it, &minus1, sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_probe, &zero, sizeof(int), hipMemcpyHostToDevice, c->stream));
    // snapshots: the run writes at most min(box * snapshots, host capacity)
    // floats, whatever an earlier run on this context allocated
    c->snap_lim = 0;
    if (c->snap_every > 0) {
        const long long W = snap_cells(c);
        const long long nsnap = (nt + c->snap_every - 1) / c->snap_every;
        const long long need = std::min(W * nsnap, c->snap_cap);
        if (need > c->d_snap_alloc) {
            c->d_snap_alloc = 0;
            rc = realloc_dev((void**)&c->d_snap, sizeof(float) * need);
            if (rc) return rc;
            c->d_snap_alloc = need;
        }
        c->snap_lim = need;
    }
    // per-run kernel arguments
    c->a_vz.nshots = nshots;
    c->a_sz.nshots = nshots;
    c->a_vz.rik = c->d_rik;
    c->a_sz.rik = c->d_rik;
    c->a_rec.rec_ip = c->d_rec_ip;
    c->a_rec.rx = c->d_rec;
    c->a_rec.ry = c->d_rec + (long long)c->total_rec * nt;
    c->a_rec.rz = c->d_rec + 2LL * c->total_rec * nt;
    c->a_rec.nt = nt;
    int maxr = 1;
    for (int s = 0; s < nshots; ++s) maxr = std::max(maxr, shots[s].nrec);
    c->g_rec = dim3((unsigned)((maxr + kThreads - 1) / kThreads), 1, 1);
    if (c->autotune && c->tuned != nshots) {
        rc = autotune_tiles(c);
        if (rc) return rc;
        c->tuned = nshots;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->begun = 1;
    if (c->use_graphs) {
        rc = capture_graph(c);
        if (rc) {
            c->use_graphs = false;  // fall back to direct launches of the same kernels
            qvti_clear_error();
        }
    }
    return QVTI_OK;
}

extern "C" int qvti_step(qvti_ctx* c, int nsteps)
{
    if (!c) return qvti_fail(QVTI_E_ARG, "NULL context");
    if (!c->begun) return qvti_fail(QVTI_E_STATE, "qvti_begin must be called first");
    if (nsteps < 0 || c->steps_done + nsteps > c->nt)
        return qvti_fail(QVTI_E_ARG, "%d steps requested, %d of %d done", nsteps, c->steps_done, c->nt);
    HIPCHK(hipSetDevice(c->device));
    const long long nprobe = 3LL * c->S * c->N;
    for (int k = 0; k < nsteps; ++k) {
        if (c->use_graphs && c->gexec && !c->timing) {
            HIPCHK(hipGraphLaunch(c->gexec, c->stream));
        } else {
            enqueue_step_with_snap(c);
        }
        c->steps_done++;
        if (c->probe_every > 0 && (c->steps_done % c->probe_every == 0 || c->steps_done == c->nt)) {
            const int e0 = c->timing ? record_event(c, KC_PROBE) : -1;
            hipLaunchKernelGGL(k_probe, dim3(1024), dim3(kThreads), 0, c->stream, (const float*)c->v, nprobe, c->d_probe);
            if (e0 >= 0) record_event(c, -1 - KC_PROBE);
        }
    }
    HIPCHK(hipGetLastError());
    return QVTI_OK;
}

static void harvest_events(qvti_ctx* c)
{
    for (int k = 0; k + 1 < c->ev_used; ++k) {
        const int cls = c->ev_class[k];
        if (cls < 0 || c->ev_class[k + 1] != -1 - cls) continue;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->ev[k], c->ev[k + 1]) == hipSuccess) {
            c->kc_ms[cls] += ms;
            c->kc_launches[cls]++;
        }
        ++k;
    }
    c->ev_used = 0;
}

extern "C" int qvti_sync(qvti_ctx* c)
{
    if (!c) return qvti_fail(QVTI_E_ARG, "NULL context");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->timing) harvest_events(c);
    int bits = 0;
    HIPCHK(hipMemcpy(&bits, c->d_probe, sizeof(int), hipMemcpyDeviceToHost));
    float lim = c->probe_limit;
    int lb;
    memcpy(&lb, &lim, sizeof lb);
    if (bits > lb) {
        float m;
        memcpy(&m, &bits, sizeof m);
        return qvti_fail(QVTI_E_UNSTABLE, "blow-up probe fired after %d steps: max|v| = %g (NaN/Inf or > %g)",
                         c->steps_done, (double)m, (double)lim);
    }
    return QVTI_OK;
}

extern "C" int qvti_steps_done(qvti_ctx* c) { return c ? c->steps_done : -1; }

extern "C" int qvti_finish(qvti_ctx* c, float* rec_vx, float* rec_vy, float* rec_vz)
{
    if (!c || !c->begun) return qvti_fail(QVTI_E_STATE, "no run in progress");
    int rc = qvti_sync(c);
    const long long n = (long long)c->total_rec * c->nt;
    float* outs[3] = {rec_vx, rec_vy, rec_vz};
    for (int k = 0; k < 3; ++k)
        if (outs[k] && n > 0)
            HIPCHK(hipMemcpy(outs[k], c->d_rec + k * n, sizeof(float) * n, hipMemcpyDeviceToHost));
    if (c->snap_every > 0 && c->d_snap && c->snap_host) {
        const long long W = snap_cells(c);
        const long long nsnap = (c->steps_done - 1) / c->snap_every + 1;
        const long long cnt = std::min(W * nsnap, c->snap_lim);
        HIPCHK(hipMemcpy(c->snap_host, c->d_snap, sizeof(float) * cnt, hipMemcpyDeviceToHost));
    }
    return rc;
}

// ------------------------------------------------------ checkpoint ------
// Image of a run in progress: header, then for shots [0, nshots) the fields
// that carry from one step to the next (v, sig, the C-PML slabs), the records
// gathered so far and the snapshots taken so far.  dv, the spectral buffers
// and the derivative buffers are rebuilt inside every step.
namespace {
struct CkptHeader {
    char magic[8];
    int version, nx, ny, nz, pml, vti, nshots, nt, steps_done, total_rec, snap_every, snap_field, probe;
    int snap_lo[3], snap_hi[3];
    long long nfloats;
};
const char kCkptMagic[8] = {'Q', 'V', 'T', 'I', 'C', 'K', 'P', '1'};

// (device pointer, floats) regions of the image, in order
std::vector<std::pair<float*, long long>> ckpt_regions(qvti_ctx* c)
{
    std::vector<std::pair<float*, long long>> r;
    const long long n = (long long)c->nshots * c->N;
    for (int f = 0; f < 3; ++f) r.push_back({fld(c->v, f, c), n});
    for (int f = 0; f < 6; ++f) r.push_back({fld(c->sig, f, c), n});
    for (int a = 0; a < 3; ++a) {
        const long long sl = (long long)c->nshots * c->slab[a], stride = (long long)c->S * c->slab[a];
        for (int f = 0; f < 3; ++f) r.push_back({c->phiv[a] + f * stride, sl});
        for (int f = 0; f < c->nphis[a]; ++f) r.push_back({c->phis[a] + f * stride, sl});
    }
    if (c->total_rec > 0) r.push_back({c->d_rec, 3LL * c->total_rec * c->nt});
    if (c->snap_every > 0 && c->d_snap && c->steps_done > 0) {
        const long long W = snap_cells(c);
        const long long taken = (c->steps_done - 1) / c->snap_every + 1;
        r.push_back({c->d_snap, std::min(W * taken, c->snap_lim)});
    }
    return r;
}

CkptHeader ckpt_header(qvti_ctx* c, int probe)
{
    CkptHeader h;
    memset(&h, 0, sizeof h);
    memcpy(h.magic, kCkptMagic, 8);
    h.version = 1;
    h.nx = c->g.nx; h.ny = c->g.ny; h.nz = c->g.nz; h.pml = c->g.pml; h.vti = c->g.vti;
    h.nshots = c->nshots; h.nt = c->nt; h.steps_done = c->steps_done; h.total_rec = c->total_rec;
    h.snap_every = c->snap_every; h.snap_field = c->snap_field; h.probe = probe;
    for (int a = 0; a < 3; ++a) { h.snap_lo[a] = c->snap_lo[a]; h.snap_hi[a] = c->snap_hi[a]; }
    for (auto& p : ckpt_regions(c)) h.nfloats += p.second;
    return h;
}
}  // namespace

extern "C" long long qvti_checkpoint_bytes(qvti_ctx* c)
{
    if (!c || !c->begun) return qvti_fail(QVTI_E_STATE, "no run in progress");
    return (long long)sizeof(CkptHeader) + 4LL * ckpt_header(c, 0).nfloats;
}

extern "C" int qvti_checkpoint(qvti_ctx* c, void* buf, long long capacity)
{
    if (!c || !buf) return qvti_fail(QVTI_E_ARG, "NULL argument");
    if (!c->begun) return qvti_fail(QVTI_E_STATE, "no run in progress");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    int probe = 0;
    HIPCHK(hipMemcpy(&probe, c->d_probe, sizeof(int), hipMemcpyDeviceToHost));
    const CkptHeader h = ckpt_header(c, probe);
    const long long need = (long long)sizeof h + 4LL * h.nfloats;
    if (capacity < need) return qvti_fail(QVTI_E_ARG, "checkpoint needs %lld bytes, buffer holds %lld", need, capacity);
    char* p = static_cast<char*>(buf);
    memcpy(p, &h, sizeof h);
    p += sizeof h;
    for (auto& r : ckpt_regions(c)) {
        HIPCHK(hipMemcpy(p, r.first, 4 * r.second, hipMemcpyDeviceToHost));
        p += 4 * r.second;
    }
    return QVTI_OK;
}

extern "C" int qvti_restore(qvti_ctx* c, const void* buf, long long size)
{
    if (!c || !buf) return qvti_fail(QVTI_E_ARG, "NULL argument");
    if (!c->begun) return qvti_fail(QVTI_E_STATE, "qvti_begin the run to restore first");
    if (c->steps_done != 0) return qvti_fail(QVTI_E_STATE, "restore before the first qvti_step");
    CkptHeader h;
    if (size < (long long)sizeof h) return qvti_fail(QVTI_E_ARG, "checkpoint image too small");
    memcpy(&h, buf, sizeof h);
    if (memcmp(h.magic, kCkptMagic, 8) != 0 || h.version != 1) return qvti_fail(QVTI_E_ARG, "not a checkpoint image");
    if (h.nx != c->g.nx || h.ny != c->g.ny || h.nz != c->g.nz || h.pml != c->g.pml || h.vti != c->g.vti ||
        h.nshots != c->nshots || h.nt != c->nt || h.total_rec != c->total_rec || h.snap_every != c->snap_every ||
        h.snap_field != c->snap_field || memcmp(h.snap_lo, c->snap_lo, sizeof h.snap_lo) != 0 ||
        memcmp(h.snap_hi, c->snap_hi, sizeof h.snap_hi) != 0)
        return qvti_fail(QVTI_E_ARG, "checkpoint of a different run (grid, shots, nt, receivers or snapshots differ)");
    if (h.steps_done < 0 || h.steps_done > h.nt) return qvti_fail(QVTI_E_ARG, "checkpoint step %d out of range", h.steps_done);
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->steps_done = h.steps_done;  // sizes the snapshot region
    const auto regions = ckpt_regions(c);
    long long nf = 0;
    for (auto& r : regions) nf += r.second;
    if (nf != h.nfloats || size < (long long)sizeof h + 4 * nf) {
        c->steps_done = 0;
        return qvti_fail(QVTI_E_ARG, "checkpoint image size does not match the run");
    }
    const char* p = static_cast<const char*>(buf) + sizeof h;
    bool ok = true;
    for (auto& r : regions) {
        ok = ok && hipMemcpy(r.first, p, 4 * r.second, hipMemcpyHostToDevice) == hipSuccess;
        p += 4 * r.second;
    }
    const int it = h.steps_done - 1;  // vel_x advances it at the start of each step
    ok = ok && hipMemcpy(c->d_it, &it, sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(c->d_probe, &h.probe, sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        // a partly loaded state must not be stepped: the run has to be begun again
        c->steps_done = 0;
        c->begun = 0;
        return qvti_fail(QVTI_E_DEVICE, "checkpoint upload failed");
    }
    return QVTI_OK;
}

extern "C" int qvti_get_state(qvti_ctx* c, int shot, int field, float* out)
{
    if (!c || !out) return qvti_fail(QVTI_E_ARG, "NULL argument");
    if (shot < 0 || shot >= c->S || field < 0 || field > 8) return qvti_fail(QVTI_E_ARG, "shot/field out of range");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    const float* base = field < 3 ? fld(c->v, field, c) : fld(c->sig, field - 3, c);
    HIPCHK(hipMemcpy(out, base + (long long)shot * c->N, sizeof(float) * c->N, hipMemcpyDeviceToHost));
    return QVTI_OK;
}

extern "C" int qvti_forward_batch(qvti_ctx* c, int nt, const float* rik, const qvti_shot* shots, int nshots,
                                  float* rec_vx, float* rec_vy, float* rec_vz)
{
    int rc = qvti_begin(c, nt, rik, shots, nshots);
    if (rc) return rc;
    // chunks so that a blow-up ends the run early
    const int chunk = 200;
    for (int done = 0; done < nt;) {
        const int k = std::min(chunk, nt - done);
        rc = qvti_step(c, k);
        if (rc) return rc;
        done += k;
        rc = qvti_sync(c);
        if (rc) break;
    }
    const int rc2 = qvti_finish(c, rec_vx, rec_vy, rec_vz);
    return rc ? rc : rc2;
}

extern "C" int qvti_forward(qvti_ctx* c, int nt, const float* rik, const qvti_shot* shot, float* rec_vx,
                            float* rec_vy, float* rec_vz)
{
    return qvti_forward_batch(c, nt, rik, shot, 1, rec_vx, rec_vy, rec_vz);
}

// ------------------------------------------------------- batch width ----
// Shots per launch for batched runs, tuned once per (grid, k-space mode,
// min(context capacity, shot count)) and kept for the process: the results
// do not depend on it (every shot's arithmetic is the same in any batch), only
// the throughput does (the medium changes coefficient values, not the work).
namespace {
struct WidthKey {
    int nx, ny, nz, vti, ks, cap;
    bool operator<(const WidthKey& o) const
    {
        return std::tie(nx, ny, nz, vti, ks, cap) < std::tie(o.nx, o.ny, o.nz, o.vti, o.ks, o.cap);
    }
};
struct WidthChoice {
    int width;
    std::vector<std::pair<int, double>> trials;  // (width, ms per shot-step)
};
std::mutex g_width_mu;
std::map<WidthKey, WidthChoice> g_width;
}  // namespace

// ms per shot-step of W shots stepped together: a silent run (zero wavelet,
// so the state stays zero) of 2 warm-up and 6 timed steps on the context.
static int time_width(qvti_ctx* c, const qvti_shot* shots, int nshots, int W, double* ms)
{
    const int warm = 2, timed = 6;
    std::vector<float> silent(warm + timed, 0.f);
    std::vector<qvti_shot> sh(W);
    for (int s = 0; s < W; ++s) sh[s] = shots[s % nshots];
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        hipEventDestroy(e0);
        return qvti_fail(QVTI_E_DEVICE, "hipEventCreate failed");
    }
    int rc = qvti_begin(c, warm + timed, silent.data(), sh.data(), W);
    if (rc) {
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        return rc;
    }
    const int probe = c->probe_every;
    c->probe_every = 0;
    rc = qvti_step(c, warm);
    if (!rc && hipEventRecord(e0, c->stream) != hipSuccess) rc = qvti_fail(QVTI_E_DEVICE, "event record failed");
    if (!rc) rc = qvti_step(c, timed);
    float t = 0.f;
    if (!rc && (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&t, e0, e1) != hipSuccess))
        rc = qvti_fail(QVTI_E_DEVICE, "width timing failed");
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    c->probe_every = probe;
    c->begun = 0;  // the timing run is not a run the caller can continue
    *ms = (double)t / timed / W;
    return rc;
}

extern "C" int qvti_batch_width(qvti_ctx* c, const qvti_shot* shots, int nshots, int* width, int max_trials,
                                int* trial_width, double* trial_ms)
{
    if (!c || !shots || !width || nshots < 1) return qvti_fail(QVTI_E_ARG, "bad arguments");
    if (!c->has_model) return qvti_fail(QVTI_E_STATE, "qvti_set_model must be called first");
    const int cap = std::min(c->S, nshots);
    const WidthKey key{c->g.nx, c->g.ny, c->g.nz, c->g.vti, c->ks ? 1 : 0, cap};
    WidthChoice choice;
    bool found = false;
    {
        std::lock_guard<std::mutex> lk(g_width_mu);
        auto it = g_width.find(key);
        if (it != g_width.end()) {
            choice = it->second;
            found = true;
        }
    }
    if (!found) {
        std::vector<int> cand;
        for (int w = 1; w < cap; w *= 2) cand.push_back(w);
        cand.push_back(cap);
        double best = 1e300;
        choice.width = 1;
        for (int w : cand) {
            double ms = 0.0;
            const int rc = time_width(c, shots, nshots, w, &ms);
            if (rc) return rc;
            choice.trials.push_back({w, ms});
            if (ms < best) {
                best = ms;
                choice.width = w;
            }
        }
        std::lock_guard<std::mutex> lk(g_width_mu);
        g_width[key] = choice;
    }
    *width = choice.width;
    const int n = std::min<int>(max_trials, (int)choice.trials.size());
    for (int k = 0; k < n; ++k) {
        if (trial_width) trial_width[k] = choice.trials[k].first;
        if (trial_ms) trial_ms[k] = choice.trials[k].second;
    }
    return (int)choice.trials.size();
}

extern "C" int qvti_forward_shots(qvti_ctx* c, int nt, const float* rik, const qvti_shot* shots, int nshots,
                                  float* rec_vx, float* rec_vy, float* rec_vz)
{
    int W = 1;
    const int rc = qvti_batch_width(c, shots, nshots, &W, 0, nullptr, nullptr);
    if (rc < 0) return rc;
    long long off = 0;  // floats of the records written so far, per component
    for (int s0 = 0; s0 < nshots; s0 += W) {
        const int k = std::min(W, nshots - s0);
        const int r = qvti_forward_batch(c, nt, rik, shots + s0, k, rec_vx ? rec_vx + off : nullptr,
                                         rec_vy ? rec_vy + off : nullptr, rec_vz ? rec_vz + off : nullptr);
        if (r) return r;
        for (int s = s0; s < s0 + k; ++s) off += (long long)nt * shots[s].nrec;
    }
    return QVTI_OK;
}

extern "C" int qvti_forward_acoustic_3D(qvti_ctx* c, int myid, int is, int nt, int ntx, int nty, int ntz, int ntp,
                                        int nx, int ny, int nz, int pml, float dx, float dy, float dz, float dt,
                                        float f0, float w0, float velp_max, const float* rik, const float* velp,
                                        const float* gama_p, const float* vels, const float* gama_s,
                                        const float* rho, const qvti_ref_source* ss, int rnmax, int rnx_max,
                                        int rny_max, int dr, float* record, float* record2, float* record3)
{
    (void)myid; (void)nx; (void)ny; (void)nz; (void)velp_max; (void)dr;
    if (!c || !ss) return qvti_fail(QVTI_E_ARG, "NULL argument");
    const qvti_grid& g = c->g;
    if (ntx != g.nx || nty != g.ny || ntz != g.nz || (long long)ntp != c->N || pml != g.pml || dx != g.dx ||
        dy != g.dy || dz != g.dz || dt != g.dt || f0 != g.f0 || w0 != g.w0)
        return qvti_fail(QVTI_E_ARG, "arguments disagree with the context's grid");
    const int R = rnx_max * rny_max;
    if (R < 0 || R > rnmax) return qvti_fail(QVTI_E_ARG, "rnx_max*rny_max exceeds rnmax");
    int rc = qvti_set_model(c, velp, gama_p, vels, gama_s, rho, nullptr, nullptr);
    if (rc) return rc;
    const qvti_ref_source& s = ss[is];
    std::vector<int> rz(R, s.r_iz);
    qvti_shot sh;
    sh.kind = QVTI_SRC_FX;
    sh.sx = s.s_ix;
    sh.sy = s.s_iy;
    sh.sz = s.s_iz + 10;  // cuviscoelas_3D.cu:772
    sh.nrec = R;
    sh.rx = s.r_ix;
    sh.ry = s.r_iy;
    sh.rz = rz.data();
    return qvti_forward(c, nt, rik, &sh, record, record2, record3);
}

extern "C" int qvti_set_snapshot(qvti_ctx* c, int field, int every, float* host_buf, long long capacity)
{
    if (!c) return qvti_fail(QVTI_E_ARG, "NULL context");
    if (field < 0 || field > 8 || every < 0 || (every > 0 && (!host_buf || capacity <= 0)))
        return qvti_fail(QVTI_E_ARG, "bad snapshot arguments");
    c->snap_field = field;
    c->snap_every = every;
    c->snap_host = host_buf;
    c->snap_cap = capacity;
    c->begun = 0;  // a new qvti_begin is needed (the step graph changes)
    return QVTI_OK;
}

extern "C" int qvti_set_snapshot_window(qvti_ctx* c, const int* lo, const int* hi)
{
    if (!c || !lo || !hi) return qvti_fail(QVTI_E_ARG, "NULL argument");
    const int n[3] = {c->g.nx, c->g.ny, c->g.nz};
    for (int a = 0; a < 3; ++a)
        if (lo[a] < 0 || lo[a] >= hi[a] || hi[a] > n[a])
            return qvti_fail(QVTI_E_ARG, "snapshot window axis %d: [%d, %d) not inside [0, %d)", a, lo[a], hi[a], n[a]);
    for (int a = 0; a < 3; ++a) {
        c->snap_lo[a] = lo[a];
        c->snap_hi[a] = hi[a];
    }
    c->begun = 0;
    return QVTI_OK;
}

extern "C" int qvti_set_media(qvti_ctx* c, int media)
{
    if (!c) return qvti_fail(QVTI_E_ARG, "NULL context");
    if (media < QVTI_MEDIA_VISCO || media > QVTI_MEDIA_QCOMP) return qvti_fail(QVTI_E_ARG, "unknown medium %d", media);
    c->media = media;
    c->has_model = false;  // the coefficients are built by the next qvti_set_model
    return QVTI_OK;
}

extern "C" int qvti_set_split(qvti_ctx* c, int axis, int variant)
{
    if (!c) return qvti_fail(QVTI_E_ARG, "NULL context");
    if (axis < 0 || axis > 2) return qvti_fail(QVTI_E_ARG, "axis %d", axis);
    const int n = axis == 0 ? c->g.nx : (axis == 1 ? c->g.ny : c->g.nz);
    const LenOps* o = len_ops(n, variant);
    if (!o) return qvti_fail(QVTI_E_ARG, "length %d has no split variant %d", n, variant);
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->begun = 0;  // the launch sequence changes: begin again
    c->split[axis] = variant;
    c->ops[axis] = o;
    int pw[KC_COUNT], jz[KC_COUNT];
    std::copy(c->pw, c->pw + KC_COUNT, pw);
    std::copy(c->jz, c->jz + KC_COUNT, jz);
    const int rc = build_launch_config(c);
    if (rc) return rc;
    if (!c->autotune) {
        std::copy(pw, pw + KC_COUNT, c->pw);
        std::copy(jz, jz + KC_COUNT, c->jz);
        apply_tiles(c);
    }
    c->tuned = 0;
    return QVTI_OK;
}

extern "C" int qvti_get_split(qvti_ctx* c, int* variants)
{
    if (!c || !variants) return qvti_fail(QVTI_E_ARG, "NULL argument");
    for (int a = 0; a < 3; ++a) variants[a] = c->split[a];
    return QVTI_OK;
}

static int time_width(qvti_ctx* c, const qvti_shot* shots, int nshots, int W, double* ms);

extern "C" int qvti_tune_split(qvti_ctx* c, int* variants, double* ms_per_step)
{
    if (!c) return qvti_fail(QVTI_E_ARG, "NULL context");
    if (!c->has_model) return qvti_fail(QVTI_E_STATE, "qvti_set_model must be called first");
    const int dims[3] = {c->g.nx, c->g.ny, c->g.nz};
    qvti_shot sh{};
    sh.kind = QVTI_SRC_FZ;
    sh.sx = dims[0] / 2;
    sh.sy = dims[1] / 2;
    sh.sz = dims[2] / 2;
    // one axis at a time: time a silent run with each variant, keep the faster
    for (int a = 0; a < 3; ++a) {
        if (!len_ops(dims[a], 1)) continue;
        int best_v = c->split[a];
        double best = 1e300;
        for (int v = 0; v < 2; ++v) {
            int rc = qvti_set_split(c, a, v);
            double ms = 0.0;
            if (!rc) rc = time_width(c, &sh, 1, 1, &ms);
            if (rc) return rc;
            if (ms_per_step) ms_per_step[2 * a + v] = ms;
            if (ms < best) {
                best = ms;
                best_v = v;
            }
        }
        const int rc = qvti_set_split(c, a, best_v);
        if (rc) return rc;
    }
    {
        std::lock_guard<std::mutex> lk(g_split_mu);
        g_split[grid_key(c->g)] = {c->split[0], c->split[1], c->split[2]};
    }
    if (variants)
        for (int a = 0; a < 3; ++a) variants[a] = c->split[a];
    return QVTI_OK;
}

extern "C" int qvti_save_wisdom(const char* path)
{
    if (!path) return qvti_fail(QVTI_E_ARG, "NULL path");
    FILE* f = fopen(path, "w");
    if (!f) return qvti_fail(QVTI_E_ARG, "cannot write %s", path);
    std::lock_guard<std::mutex> lk(g_split_mu);
    fprintf(f, "qvti-wisdom 1\n");
    for (const auto& e : g_split)
        fprintf(f, "%d %d %d %d %d %d %d\n", e.first[0], e.first[1], e.first[2], e.first[3], e.second[0], e.second[1],
                e.second[2]);
    const bool ok = fclose(f) == 0;
    return ok ? QVTI_OK : qvti_fail(QVTI_E_ARG, "writing %s failed", path);
}

extern "C" int qvti_load_wisdom(const char* path)
{
    if (!path) return qvti_fail(QVTI_E_ARG, "NULL path");
    FILE* f = fopen(path, "r");
    if (!f) return qvti_fail(QVTI_E_ARG, "cannot read %s", path);
    char head[32] = {};
    int version = 0;
    if (fscanf(f, "%31s %d", head, &version) != 2 || strcmp(head, "qvti-wisdom") != 0 || version != 1) {
        fclose(f);
        return qvti_fail(QVTI_E_ARG, "%s is not a qvti wisdom file", path);
    }
    std::map<GridKey, std::array<int, 3>> add;
    int v[7];
    int got;
    while ((got = fscanf(f, "%d %d %d %d %d %d %d", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6])) == 7) {
        for (int a = 0; a < 3; ++a)
            if (!len_ops(v[a], 0) || !len_ops(v[a], v[4 + a])) {
                fclose(f);
                return qvti_fail(QVTI_E_ARG, "%s: no split variant %d for length %d", path, v[4 + a], v[a]);
            }
        add[GridKey{v[0], v[1], v[2], v[3] ? 1 : 0}] = {v[4], v[5], v[6]};
    }
    const bool clean = got == EOF;
    fclose(f);
    if (!clean) return qvti_fail(QVTI_E_ARG, "%s: malformed line", path);
    std::lock_guard<std::mutex> lk(g_split_mu);
    for (const auto& e : add) g_split[e.first] = e.second;
    return QVTI_OK;
}

extern "C" int qvti_set_graphs(qvti_ctx* c, int on)
{
    if (!c) return qvti_fail(QVTI_E_ARG, "NULL context");
    c->use_graphs = on != 0;
    return QVTI_OK;
}

extern "C" int qvti_set_autotune(qvti_ctx* c, int mode)
{
    if (!c) return qvti_fail(QVTI_E_ARG, "NULL context");
    const bool split = (mode & 16) != 0;
    const int m = mode & ~16;
    if ((m != 0 && m != 1 && m != 2 && m != 4) || (m == 1 && split))
        return qvti_fail(QVTI_E_ARG, "autotune mode %d not in {0, 1, 2, 4, 16, 18, 20}", mode);
    c->autotune = m == 1;
    c->tuned = 0;
    const int w = m == 0 ? 1 : m;
    if (!c->autotune)
        for (int k = 0; k < KC_COUNT; ++k) {
            c->pw[k] = w;  // the widest available <= w is used
            c->jz[k] = split ? class_fields(c, k) : 1;
        }
    apply_tiles(c);
    return QVTI_OK;
}

extern "C" int qvti_set_kspace(qvti_ctx* c, float vref, float fcut, float vmax, float taper_ratio)
{
    if (!c) return qvti_fail(QVTI_E_ARG, "NULL context");
    if (!(vref >= 0.f) || !std::isfinite(vref)) return qvti_fail(QVTI_E_ARG, "vref %g must be >= 0", vref);
    if (!(fcut >= 0.f) || !std::isfinite(fcut)) return qvti_fail(QVTI_E_ARG, "fcut %g must be >= 0", fcut);
    if (fcut > 0.f) {
        if (!(vmax > 0.f) || !std::isfinite(vmax)) return qvti_fail(QVTI_E_ARG, "vmax %g must be > 0", vmax);
        if (!(taper_ratio >= 0.f && taper_ratio <= 1.f))
            return qvti_fail(QVTI_E_ARG, "taper_ratio %g not in [0, 1]", taper_ratio);
    }
    HIPCHK(hipSetDevice(c->device));
    // the launch arguments change: let queued steps finish, and require a
    // new qvti_begin (as qvti_set_snapshot does)
    HIPCHK(hipStreamSynchronize(c->stream));
    c->begun = 0;
    const bool on = vref > 0.f || fcut > 0.f;
    if (on && !c->H2) {
        const int rc = dalloc(c, (void**)&c->H2, sizeof(float) * 6 * c->S * c->N);
        if (rc) return rc;
    }
    c->ks = on;
    c->ks_vref = vref;
    c->ks_fcut = fcut;
    c->ks_vmax = fcut > 0.f ? vmax : 0.f;
    c->ks_taper = fcut > 0.f ? taper_ratio : 0.f;
    // the launch arguments and the y-inverse field count change: rebuild
    // them, keep the fixed tile widths a non-autotuned context was given
    int pw[KC_COUNT], jz[KC_COUNT];
    std::copy(c->pw, c->pw + KC_COUNT, pw);
    std::copy(c->jz, c->jz + KC_COUNT, jz);
    const int rc = build_launch_config(c);
    if (rc) return rc;
    if (!c->autotune) {
        std::copy(pw, pw + KC_COUNT, c->pw);
        for (int k = 0; k < KC_COUNT; ++k) c->jz[k] = jz[k] > 1 ? class_fields(c, k) : 1;
        apply_tiles(c);
    }
    c->tuned = 0;
    return QVTI_OK;
}

extern "C" int qvti_tile_choice(qvti_ctx* c, int max_classes, int* pencils, int* split)
{
    if (!c || !pencils) return qvti_fail(QVTI_E_ARG, "NULL argument");
    const int n = std::min(max_classes, (int)KC_COUNT);
    for (int k = 0; k < n; ++k) {
        const LenOps* o = nullptr;
        if (k == KC_VEL_X || k == KC_XPASS || k == KC_STR_X) o = c->ops[0];
        if (k == KC_VEL_Y || k == KC_Y_FWD || k == KC_Y_INV || k == KC_STR_Y) o = c->ops[1];
        pencils[k] = o ? tile_w(o, c->pw[k]).P : 0;
        if (split) split[k] = o ? c->jz[k] : 0;
    }
    return n;
}

extern "C" int qvti_set_probe(qvti_ctx* c, int every, float limit)
{
    if (!c || every < 0 || !(limit > 0)) return qvti_fail(QVTI_E_ARG, "bad probe arguments");
    c->probe_every = every;
    c->probe_limit = limit;
    return QVTI_OK;
}

extern "C" int qvti_enable_kernel_timing(qvti_ctx* c, int max_steps)
{
    if (!c) return qvti_fail(QVTI_E_ARG, "NULL context");
    HIPCHK(hipSetDevice(c->device));
    for (auto e : c->ev) hipEventDestroy(e);
    c->ev.clear();
    c->ev_class.clear();
    c->ev_used = 0;
    for (int k = 0; k < KC_COUNT; ++k) { c->kc_ms[k] = 0; c->kc_launches[k] = 0; }
    c->timing = max_steps > 0;
    if (!c->timing) return QVTI_OK;
    c->ev_max = max_steps * 2 * (KC_COUNT + 1);
    c->ev.resize(c->ev_max);
    c->ev_class.resize(c->ev_max);
    for (int k = 0; k < c->ev_max; ++k) HIPCHK(hipEventCreate(&c->ev[k]));
    return QVTI_OK;
}

extern "C" int qvti_kernel_times(qvti_ctx* c, int max_classes, char* names, int name_len, long long* launches,
                                 double* total_ms)
{
    if (!c) return qvti_fail(QVTI_E_ARG, "NULL context");
    const int n = std::min(max_classes, (int)KC_COUNT);
    for (int k = 0; k < n; ++k) {
        if (names && name_len > 0) {
            strncpy(names + (long long)k * name_len, kClassNames[k], name_len - 1);
            names[(long long)k * name_len + name_len - 1] = 0;
        }
        if (launches) launches[k] = c->kc_launches[k];
        if (total_ms) total_ms[k] = c->kc_ms[k];
    }
    return n;
}

// Compulsory HBM bytes of one launch per kernel class: every input field read
// once, every output written once, over all shots of the launch.
extern "C" int qvti_kernel_bytes(qvti_ctx* c, int max_classes, double* bytes)
{
    if (!c || !bytes) return qvti_fail(QVTI_E_ARG, "NULL argument");
    const double N = (double)c->N * std::max(1, c->nshots), f = 4.0 * N;
    const double slab = (double)std::max(1, c->nshots) * 4.0;
    const double sx = c->slab[0] * slab, sy = c->slab[1] * slab, sz = c->slab[2] * slab;
    const int ncoef = c->g.vti ? 25 : 15;
    double b[KC_COUNT] = {};
    b[KC_TICK] = 8;
    b[KC_VEL_X] = 6 * f + 3 * 2 * sx;
    b[KC_VEL_Y] = 9 * f + 3 * 2 * sy;
    b[KC_VEL_Z] = (3 + 3 + 3 + 3 + 3 + 6) * f + 3.0 * 4 * c->N + 3 * 2 * sz;  // sig, acc, v r/w, dv, P, buoy
    b[KC_Y_FWD] = 12 * f;
    // k-space mode: three more complex spectra through the x-pass, y-inverse
    // and z-inverse, six more real fields out
    b[KC_XPASS] = 6 * f + (c->ks ? 16 : 10) * f;
    b[KC_Y_INV] = (c->ks ? 32 : 20) * f;
    b[KC_Z_INV] = c->ks ? 16 * f + 15 * f : 10 * f + 9 * f;
    b[KC_STR_X] = 30 * f;
    b[KC_STR_Y] = 30 * f;
    b[KC_STR_Z] = (15 + 30 + 12) * f + ncoef * 4.0 * c->N + 2 * (5 * sx + 5 * sy + 4 * sz);
    b[KC_RECORD] = 6.0 * 4 * c->total_rec;
    const int n = std::min(max_classes, (int)KC_COUNT);
    for (int k = 0; k < n; ++k) bytes[k] = b[k];
    return n;
}

// ------------------------------------------------ standalone transforms ---
// W_n^i for the standalone transforms (twiddle_w: the same bits as the
// device-built tables of qvti_fft3d_noalloc)
static std::vector<float2> twiddles(int n)
{
    std::vector<float2> t(n);
    for (int i = 0; i < n; ++i) t[i] = twiddle_w(i, n);
    return t;
}

// In-place unnormalised 3-D C2C of d [nx][ny][nz] on stream st: z rows,
// then the y and x axes as strided pencils; tw[a] holds W_{n_a}^e.
static void enqueue_fft3d(int nx, int ny, int nz, int sign, float2* d, const float2* const tw[3], hipStream_t st)
{
    const int dims[3] = {nx, ny, nz};
    const long long N = (long long)nx * ny * nz, nrows = (long long)nx * ny;
    const LenOps& Z = *len_ops(nz);
    ZRowsArgs z{};
    z.nrows = nrows; z.shot_stride = N; z.tw = tw[2]; z.mode = 2; z.dir = sign; z.nfields = 1;
    z.cin[0] = d; z.cout[0] = d;
    Z.zrows(dim3((unsigned)((nrows + Z.Prow - 1) / Z.Prow), 1), st, z);
    for (int axis = 1; axis >= 0; --axis) {
        const LenOps& A = *len_ops(dims[axis]);
        CStridedArgs y{};
        y.nz = nz; y.sx = (long long)ny * nz; y.sy = nz; y.axis = axis;
        y.ztiles = (nz + A.P[0] - 1) / A.P[0]; y.shot_stride = N; y.tw = tw[axis]; y.nfields = 1; y.f[0] = d;
        const int other = axis ? nx : ny;
        A.cfft[sign < 0 ? 0 : 1][0](dim3((unsigned)(other * y.ztiles), 1), st, y);
    }
}

extern "C" int qvti_fft3d(int nx, int ny, int nz, int sign, float* host_data)
{
    if (!host_data || (sign != 1 && sign != -1)) return qvti_fail(QVTI_E_ARG, "bad arguments");
    const int dims[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a)
        if (!supported_n(dims[a]))
            return qvti_fail(QVTI_E_UNSUPPORTED, "axis %d length %d: supported lengths are %s", a, dims[a], supported_lengths());
    const long long N = (long long)nx * ny * nz;
    float2* d = nullptr;
    float2* tw[3] = {nullptr, nullptr, nullptr};
    HIPCHK(hipMalloc(&d, sizeof(float2) * N));
    int rc = QVTI_OK;
    for (int a = 0; a < 3 && !rc; ++a) {
        std::vector<float2> t = twiddles(dims[a]);
        if (hipMalloc(&tw[a], sizeof(float2) * dims[a]) != hipSuccess ||
            hipMemcpy(tw[a], t.data(), sizeof(float2) * dims[a], hipMemcpyHostToDevice) != hipSuccess)
            rc = qvti_fail(QVTI_E_DEVICE, "table upload failed: %s", hipGetErrorString(hipGetLastError()));
    }
    if (!rc && hipMemcpy(d, host_data, sizeof(float2) * N, hipMemcpyHostToDevice) != hipSuccess)
        rc = qvti_fail(QVTI_E_DEVICE, "H2D failed");
    if (!rc) {
        const float2* const tws[3] = {tw[0], tw[1], tw[2]};
        enqueue_fft3d(nx, ny, nz, sign, d, tws, nullptr);
        if (hipDeviceSynchronize() != hipSuccess || hipGetLastError() != hipSuccess)
            rc = qvti_fail(QVTI_E_DEVICE, "fft kernels failed");
    }
    if (!rc && hipMemcpy(host_data, d, sizeof(float2) * N, hipMemcpyDeviceToHost) != hipSuccess)
        rc = qvti_fail(QVTI_E_DEVICE, "D2H failed");
    hipFree(d);
    for (auto p : tw) if (p) hipFree(p);
    return rc;
}

// ------------------------------------------- allocation-free transform ---
// qvti_fft3d_noalloc keeps its only device memory beyond the caller's data -
// the three axes' twiddle tables W_n^e, e < n <= kMaxN - in this file-scope
// array.  Each GPU holds its own copy; host code reaches it only through the
// per-device address from hipGetSymbolAddress, never through the symbol.
constexpr int kNoAllocTw = 3 * kMaxN;  // float2 entries: 12 KiB per device
static __device__ __align__(256) float2 scratch[kNoAllocTw];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float2* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once per
// device, with that device current).
int scratch_ptr(int dev, float2** out)
{
    if (dev < 0 || dev >= kMaxDevices) return qvti_fail(QVTI_E_DEVICE, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return qvti_fail(QVTI_E_DEVICE, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<float2*>(p);
    }
    *out = g_scratch_ptr[dev];
    return QVTI_OK;
}
}  // namespace

// tw[a * kMaxN + e] = W_{n_a}^e for e < n_a (the three axes' tables)
__global__ __launch_bounds__(256) void k_twiddle_tables(float2* tw, int nx, int ny, int nz)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int a = t / kMaxN, e = t - a * kMaxN;
    const int n = a == 0 ? nx : (a == 1 ? ny : nz);
    if (a < 3 && e < n) tw[t] = twiddle_w(e, n);
}

extern "C" int qvti_fft3d_noalloc(int nx, int ny, int nz, int sign, float* d_data, void* stream)
{
    if (!d_data || (sign != 1 && sign != -1)) return qvti_fail(QVTI_E_ARG, "bad arguments");
    const int dims[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a) {
        if (dims[a] > kMaxN)
            return qvti_fail(QVTI_E_UNSUPPORTED, "axis %d length %d exceeds the %d-entry twiddle scratch", a,
                             dims[a], kMaxN);
        if (!supported_n(dims[a]))
            return qvti_fail(QVTI_E_UNSUPPORTED, "axis %d length %d: supported lengths are %s", a, dims[a], supported_lengths());
    }
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    float2* tw = nullptr;
    const int rc = scratch_ptr(dev, &tw);
    if (rc) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    float2* d = reinterpret_cast<float2*>(d_data);
    // every call rebuilds the tables it uses: the array keeps the last call's
    // contents, and calls that overlap on one device race on it
    hipLaunchKernelGGL(k_twiddle_tables, dim3(3 * kMaxN / 256), dim3(256), 0, st, tw, nx, ny, nz);
    const float2* const tws[3] = {tw, tw + kMaxN, tw + 2 * kMaxN};
    enqueue_fft3d(nx, ny, nz, sign, d, tws, st);
    HIPCHK(hipGetLastError());
    return QVTI_OK;
}

extern "C" int qvti_deriv1d(int nx, int ny, int nz, float dd, int axis, int sym, const float* in, float* out)
{
    if (!in || !out || axis < 0 || axis > 2 || sym < 0 || sym > 1 || !(dd > 0))
        return qvti_fail(QVTI_E_ARG, "bad arguments");
    const int dims[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a)
        if (!supported_n(dims[a]))
            return qvti_fail(QVTI_E_UNSUPPORTED, "axis %d length %d: supported lengths are %s", a, dims[a], supported_lengths());
    const int n = dims[axis];
    std::vector<float2> tw = twiddles(n), dp, dm;
    std::vector<float> kf;
    axis_symbols(n, dd, dp, dm, kf);
    const long long N = (long long)nx * ny * nz;
    float *din = nullptr, *dout = nullptr;
    float2 *dtw = nullptr, *dsym = nullptr;
    int rc = QVTI_OK;
    if (hipMalloc(&din, sizeof(float) * N) != hipSuccess || hipMalloc(&dout, sizeof(float) * N) != hipSuccess ||
        hipMalloc(&dtw, sizeof(float2) * n) != hipSuccess || hipMalloc(&dsym, sizeof(float2) * n) != hipSuccess)
        rc = qvti_fail(QVTI_E_NOMEM, "allocation failed: %s", hipGetErrorString(hipGetLastError()));
    if (!rc && (hipMemcpy(din, in, sizeof(float) * N, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(dtw, tw.data(), sizeof(float2) * n, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(dsym, (sym ? dm : dp).data(), sizeof(float2) * n, hipMemcpyHostToDevice) != hipSuccess))
        rc = qvti_fail(QVTI_E_DEVICE, "H2D failed");
    if (!rc) {
        const LenOps& A = *len_ops(n);
        const int P = A.P[0];
        if (axis < 2) {
            StridedArgs a{};
            a.nx = nx; a.ny = ny; a.nz = nz; a.pml = 0;
            a.sx = (long long)ny * nz; a.sy = nz; a.axis = axis;
            a.ztiles = (nz + 2 * P - 1) / (2 * P);
            a.mode = MODE_STORE; a.dt = 0.f; a.shot_stride = N; a.slab_stride = 0;
            a.tw = dtw; a.sym[0] = dsym; a.sym[1] = dsym;
            a.njobs = 1;
            a.job[0].in = din; a.job[0].out = dout; a.job[0].acc = nullptr; a.job[0].phi = nullptr;
            a.job[0].sym = 0; a.job[0].half = 0;
            const int other = axis == 0 ? ny : nx;
            A.deriv[MODE_STORE][0](dim3((unsigned)(other * a.ztiles), 1), nullptr, a);
        } else {
            RowDerivArgs r{};
            const long long nrows = (long long)nx * ny;
            r.nrows = nrows; r.tw = dtw; r.sym = dsym; r.in = din; r.out = dout;
            A.deriv_rows(dim3((unsigned)((nrows + 2 * A.Prow - 1) / (2 * A.Prow))), nullptr, r);
        }
        if (hipDeviceSynchronize() != hipSuccess || hipGetLastError() != hipSuccess)
            rc = qvti_fail(QVTI_E_DEVICE, "derivative kernel failed");
    }
    if (!rc && hipMemcpy(out, dout, sizeof(float) * N, hipMemcpyDeviceToHost) != hipSuccess)
        rc = qvti_fail(QVTI_E_DEVICE, "D2H failed");
    hipFree(din); hipFree(dout); hipFree(dtw); hipFree(dsym);
    return rc;
}
