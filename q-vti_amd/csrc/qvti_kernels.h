// qvti_kernels.h -- device kernels of the MI355X time-stepping engine.
//
// Per step (reference loop: cuviscoelas_3D.cu:1255-1605):
//   velocity half-step   k_deriv_strided(x) -> k_deriv_strided(y) -> k_vel_z
//   stress half-step     k_zrows(fwd) -> k_cfft_strided(y fwd) -> k_xpass
//                        -> k_cfft_strided(y inv) -> k_zrows(inv)
//                        -> k_deriv_strided(x) -> k_deriv_strided(y) -> k_stress_z
//   source / receivers   (inside k_stress_z) -> k_record
//
// Algebra used (exact in exact arithmetic; only rounding differs from the
// reference's C2C + `.x / ntp` formulation):
//  * every spectral operator of the scheme is Hermitian (real |k|^p times a
//    1-D shift-derivative whose Nyquist symbol is taken as its real part,
//    which is what `.x` of the C2C result keeps), so two real fields packed
//    as a + ib go through ONE complex transform and come back as op(a) + i op(b);
//  * an operator acting along one axis only (all nine velocity-step
//    derivatives and the A2 = 1 stress terms) is a 1-D transform along that
//    axis: the other two axes' forward and inverse transforms cancel;
//  * A_j D_a V = D_a (A_j V): the 45 stress terms are 1-D derivatives of the
//    15 fields {A1 v, v, A3 v, A1 dv, dv} per component, with dv = v - v_prev
//    in real space (v_prev: the previous step's pre-injection velocity) and
//    the 1/dt of the time difference folded into the tau coefficients, so
//    only A1(v + i dv) and A3 v need the full 3-D transform.  v and dv are
//    packed together because their magnitudes are comparable (|dv| ~ w dt |v|);
//    packing v with dv/dt instead lets the FP32 rounding of the larger part
//    swamp the smaller one.
#pragma once
#include "qvti_fft.h"

namespace qvti {

constexpr int kMaxN = 512;  // longest axis the LDS tiles are sized for

enum { MODE_STORE = 0, MODE_VEL_FIRST = 1, MODE_VEL_ACC = 2 };

struct DJob {
    const float* in;
    float* out;
    const float* acc;
    float* phi;
    int sym;   // 0: D+, 1: D-
    int half;  // 1: half-cell averaged C-PML a, b
};

// 1-D shift-derivatives of real fields along a strided axis (x or y).
// Tile: Pc complex pencils = 2*Pc adjacent real z-columns at one value of the
// other in-plane index, loaded as float2 (two real pencils per complex one).
struct StridedArgs {
    int nx, ny, nz, pml;
    long long sx, sy;
    int axis;  // 0: x (stride sx), 1: y (stride sy)
    int n, Pc, ztiles;
    int mode;
    float dt;
    long long shot_stride, slab_stride;
    FftPlan plan;
    const float2* tw;
    const float2* sym[2];
    const float *pa, *pb, *pah, *pbh;
    int njobs;
    DJob job[15];
};

__global__ __launch_bounds__(kThreads) void k_deriv_strided(StridedArgs a)
{
    __shared__ float2 tile[kTileMax + 64];
    __shared__ float2 tws[kMaxN];
    const int shot = blockIdx.y;
    const int other = blockIdx.x / a.ztiles;
    const int z0 = (blockIdx.x - other * a.ztiles) * 2 * a.Pc;
    const int n = a.n, Pc = a.Pc, ps = n + 1, T = n * Pc, nz = a.nz;
    const long long astride = a.axis == 0 ? a.sx : a.sy;
    const long long obase = a.axis == 0 ? (long long)other * a.sy : (long long)other * a.sx;
    const long long soff = (long long)shot * a.shot_stride;
    for (int t = threadIdx.x; t < n; t += kThreads) tws[t] = a.tw[t];
    for (int jb = 0; jb < a.njobs; ++jb) {
        const DJob J = a.job[jb];
        const float* in = J.in + soff + obase;
        __syncthreads();
        for (int e = threadIdx.x; e < T; e += kThreads) {
            const int i = e / Pc, c = e - i * Pc, zc = z0 + 2 * c;
            float2 v = make_float2(0.f, 0.f);
            if (zc < nz) v = *reinterpret_cast<const float2*>(in + i * astride + zc);
            tile[c * ps + i] = v;
        }
        __syncthreads();
        fft_tile(tile, tws, a.plan, Pc, ps, -1.f);
        const float2* S = a.sym[J.sym];
        for (int e = threadIdx.x; e < T; e += kThreads) {
            const int c = e / n, i = e - c * n;
            tile[c * ps + i] = cmul(tile[c * ps + i], S[i]);
        }
        __syncthreads();
        fft_tile(tile, tws, a.plan, Pc, ps, 1.f);
        for (int e = threadIdx.x; e < T; e += kThreads) {
            const int i = e / Pc, c = e - i * Pc, zc = z0 + 2 * c;
            if (zc >= nz) continue;
            float2 d = tile[c * ps + i];
            const long long g = soff + obase + i * astride + zc;
            if (a.mode != MODE_STORE) {
                const int ix = a.axis == 0 ? i : other;
                const int iy = a.axis == 0 ? other : i;
                const int pml = a.pml;
                if ((i < pml || i >= n - pml) && ix < a.nx - 1 && iy < a.ny - 1) {
                    const int si = i < pml ? i : i - (n - 2 * pml);
                    const long long sidx = a.axis == 0 ? ((long long)si * a.ny + iy) * nz + zc
                                                       : ((long long)ix * 2 * pml + si) * nz + zc;
                    float2* ph = reinterpret_cast<float2*>(J.phi + (long long)shot * a.slab_stride + sidx);
                    const float A = J.half ? a.pah[i] : a.pa[i];
                    const float B = J.half ? a.pbh[i] : a.pb[i];
                    float2 p = *ph;
                    if (zc < nz - 1) { p.x = p.x + a.dt * (-A * p.x - B * d.x); d.x = d.x + p.x; }
                    if (zc + 1 < nz - 1) { p.y = p.y + a.dt * (-A * p.y - B * d.y); d.y = d.y + p.y; }
                    *ph = p;
                }
                if (a.mode == MODE_VEL_ACC) {
                    const float2 ac = *reinterpret_cast<const float2*>(J.acc + g);
                    d = make_float2(ac.x + d.x, ac.y + d.y);
                }
            }
            *reinterpret_cast<float2*>(J.out + g) = d;
        }
    }
}

// Load 2*Pc real rows (z-pencils) r0.. into Pc complex pencils of `tile`.
__device__ __forceinline__ void load_row_pairs(float2* tile, const float* in, long long r0, long long nrows,
                                               int n, int Pc, int ps)
{
    float* tf = reinterpret_cast<float*>(tile);
    const int T = 2 * Pc * n;
    for (int e = threadIdx.x; e < T; e += kThreads) {
        const int r = e / n, i = e - r * n;
        const long long row = r0 + r;
        tf[2 * ((r >> 1) * ps + i) + (r & 1)] = row < nrows ? in[row * n + i] : 0.f;
    }
}

// 1-D shift-derivative along z of real fields (standalone entry point).
struct RowDerivArgs {
    int nz, Pc;
    long long nrows;
    FftPlan plan;
    const float2* tw;
    const float2* sym;
    const float* in;
    float* out;
};

__global__ __launch_bounds__(kThreads) void k_deriv_rows(RowDerivArgs a)
{
    __shared__ float2 tile[kTileMax + 64];
    __shared__ float2 tws[kMaxN];
    const int n = a.nz, Pc = a.Pc, ps = n + 1, T = 2 * Pc * n;
    const long long r0 = (long long)blockIdx.x * 2 * Pc;
    for (int t = threadIdx.x; t < n; t += kThreads) tws[t] = a.tw[t];
    load_row_pairs(tile, a.in, r0, a.nrows, n, Pc, ps);
    __syncthreads();
    fft_tile(tile, tws, a.plan, Pc, ps, -1.f);
    for (int e = threadIdx.x; e < Pc * n; e += kThreads) {
        const int c = e / n, i = e - c * n;
        tile[c * ps + i] = cmul(tile[c * ps + i], a.sym[i]);
    }
    __syncthreads();
    fft_tile(tile, tws, a.plan, Pc, ps, 1.f);
    const float* tf = reinterpret_cast<const float*>(tile);
    for (int e = threadIdx.x; e < T; e += kThreads) {
        const int r = e / n, i = e - r * n;
        const long long row = r0 + r;
        if (row < a.nrows) a.out[row * n + i] = tf[2 * ((r >> 1) * ps + i) + (r & 1)];
    }
}

// Velocity update (cuviscoelas_3D.cu:1018-1101) fused with the z-derivatives:
// v += b * (acc_xy + d_z + phi_z); dv = v_new - v_old + src_prev, the change
// since the previous step's pre-injection velocity (the reference's
// V - V_prev, cuviscoelas_3D.cu:517-522, 1124-1131, without the 1/dt).
struct VelZArgs {
    int nx, ny, nz, pml;
    int Pc;
    long long nrows;
    float dt;
    long long shot_stride, slab_stride;
    FftPlan plan;
    const float2* tw;
    const float2* sym[2];
    const float *pa, *pb, *pah, *pbh;
    const float* sig[3];
    int symsel[3], half[3];
    const float* acc[3];
    float* v[3];
    float* dv[3];
    const float* buoy[3];
    float* phi[3];
    const int* src_kind;      // per shot
    const long long* src_ip;  // per shot
    const float* rik;
    const int* it;            // device step counter
};

__global__ __launch_bounds__(kThreads) void k_vel_z(VelZArgs a)
{
    __shared__ float2 tile[kTileMax + 64];
    __shared__ float2 tws[kMaxN];
    const int shot = blockIdx.y;
    const int n = a.nz, Pc = a.Pc, ps = n + 1, T = 2 * Pc * n;
    const long long r0 = (long long)blockIdx.x * 2 * Pc;
    const long long soff = (long long)shot * a.shot_stride;
    const int it = *a.it;
    const int kind = a.src_kind[shot];
    const long long sip = a.src_ip[shot];
    const float sprev = it > 0 ? a.rik[it - 1] : 0.f;
    for (int t = threadIdx.x; t < n; t += kThreads) tws[t] = a.tw[t];
    const float* tf = reinterpret_cast<const float*>(tile);
    for (int f = 0; f < 3; ++f) {
        __syncthreads();
        load_row_pairs(tile, a.sig[f] + soff, r0, a.nrows, n, Pc, ps);
        __syncthreads();
        fft_tile(tile, tws, a.plan, Pc, ps, -1.f);
        const float2* S = a.sym[a.symsel[f]];
        for (int e = threadIdx.x; e < Pc * n; e += kThreads) {
            const int c = e / n, i = e - c * n;
            tile[c * ps + i] = cmul(tile[c * ps + i], S[i]);
        }
        __syncthreads();
        fft_tile(tile, tws, a.plan, Pc, ps, 1.f);
        for (int e = threadIdx.x; e < T; e += kThreads) {
            const int r = e / n, iz = e - r * n;
            const long long row = r0 + r;
            if (row >= a.nrows) continue;
            const int ix = (int)(row / a.ny), iy = (int)(row - (long long)ix * a.ny);
            const long long ip = row * n + iz;
            float d = tf[2 * ((r >> 1) * ps + iz) + (r & 1)];
            const float vold = a.v[f][soff + ip];
            float vnew = vold;
            if (ix < a.nx - 1 && iy < a.ny - 1 && iz < n - 1) {
                if (iz < a.pml || iz >= n - a.pml) {
                    const int si = iz < a.pml ? iz : iz - (n - 2 * a.pml);
                    float* ph = a.phi[f] + (long long)shot * a.slab_stride + row * 2 * a.pml + si;
                    const float A = a.half[f] ? a.pah[iz] : a.pa[iz];
                    const float B = a.half[f] ? a.pbh[iz] : a.pb[iz];
                    const float p = *ph + a.dt * (-A * *ph - B * d);
                    *ph = p;
                    d = d + p;
                }
                const float tot = a.acc[f][soff + ip] + d;
                vnew = vold + a.buoy[f][ip] * tot;
                a.v[f][soff + ip] = vnew;
            }
            float dv = vnew - vold;
            if (ip == sip && kind == f) dv += sprev;
            a.dv[f][soff + ip] = dv;
        }
    }
}

// z-direction complex transforms on rows.
//  mode 0 (forward): P_f = FFT_z(re_in_f + i im_in_f)           (stress step start)
//  mode 1 (inverse): re_out_f + i im_out_f = IFFT_z(Q_f)         (im_out may be null)
//  mode 2 (plain):   cout_f = FFT_z^sign(cin_f)                 (standalone transforms)
struct ZRowsArgs {
    int nz, Pc;
    float sign;
    long long nrows;
    long long shot_stride;  // cells (real) == complex elements per shot
    FftPlan plan;
    const float2* tw;
    int mode, nfields;
    const float* re_in[3];
    const float* im_in[3];
    float2* cout[3];
    const float2* cin[6];
    float* re_out[6];
    float* im_out[6];
};

__global__ __launch_bounds__(kThreads) void k_zrows(ZRowsArgs a)
{
    __shared__ float2 tile[kTileMax + 64];
    __shared__ float2 tws[kMaxN];
    const int shot = blockIdx.y;
    const int n = a.nz, Pc = a.Pc, ps = n + 1, T = Pc * n;
    const long long r0 = (long long)blockIdx.x * Pc;
    const long long soff = (long long)shot * a.shot_stride;
    for (int t = threadIdx.x; t < n; t += kThreads) tws[t] = a.tw[t];
    for (int f = 0; f < a.nfields; ++f) {
        __syncthreads();
        for (int e = threadIdx.x; e < T; e += kThreads) {
            const int c = e / n, i = e - c * n;
            const long long row = r0 + c;
            float2 v = make_float2(0.f, 0.f);
            if (row < a.nrows) {
                const long long g = soff + row * n + i;
                if (a.mode == 0) v = make_float2(a.re_in[f][g], a.im_in[f][g]);
                else v = a.cin[f][g];  // modes 1 and 2
            }
            tile[c * ps + i] = v;
        }
        __syncthreads();
        fft_tile(tile, tws, a.plan, Pc, ps, a.mode == 0 ? -1.f : (a.mode == 1 ? 1.f : a.sign));
        for (int e = threadIdx.x; e < T; e += kThreads) {
            const int c = e / n, i = e - c * n;
            const long long row = r0 + c;
            if (row >= a.nrows) continue;
            const long long g = soff + row * n + i;
            const float2 v = tile[c * ps + i];
            if (a.mode == 0 || a.mode == 2) {
                a.cout[f][g] = v;
            } else {
                a.re_out[f][g] = v.x;
                if (a.im_out[f]) a.im_out[f][g] = v.y;
            }
        }
    }
}

// In-place complex transforms along a strided axis (y: stride sy, x: stride sx)
// for up to 6 complex fields.  Tile: Pc adjacent complex z-columns.
struct CStridedArgs {
    int nz;
    long long sx, sy;
    int axis, n, Pc, ztiles;
    float sign;
    long long shot_stride;
    FftPlan plan;
    const float2* tw;
    int nfields;
    float2* f[6];
};

__global__ __launch_bounds__(kThreads) void k_cfft_strided(CStridedArgs a)
{
    __shared__ float2 tile[kTileMax + 64];
    __shared__ float2 tws[kMaxN];
    const int shot = blockIdx.y;
    const int other = blockIdx.x / a.ztiles;
    const int z0 = (blockIdx.x - other * a.ztiles) * a.Pc;
    const int n = a.n, Pc = a.Pc, ps = n + 1, T = n * Pc, nz = a.nz;
    const long long astride = a.axis == 0 ? a.sx : a.sy;
    const long long obase = a.axis == 0 ? (long long)other * a.sy : (long long)other * a.sx;
    const long long soff = (long long)shot * a.shot_stride;
    for (int t = threadIdx.x; t < n; t += kThreads) tws[t] = a.tw[t];
    for (int f = 0; f < a.nfields; ++f) {
        float2* base = a.f[f] + soff + obase;
        __syncthreads();
        for (int e = threadIdx.x; e < T; e += kThreads) {
            const int i = e / Pc, c = e - i * Pc, zc = z0 + c;
            tile[c * ps + i] = zc < nz ? base[i * astride + zc] : make_float2(0.f, 0.f);
        }
        __syncthreads();
        fft_tile(tile, tws, a.plan, Pc, ps, a.sign);
        for (int e = threadIdx.x; e < T; e += kThreads) {
            const int i = e / Pc, c = e - i * Pc, zc = z0 + c;
            if (zc < nz) base[i * astride + zc] = tile[c * ps + i];
        }
    }
}

// x-pass of the 3-D pipeline: forward x-FFT of P (already z- and y-transformed),
// multiply by |k|^-1 / N and |k| / N (|k|^2 = kx^2+ky^2+kz^2 from the folded
// 1-D |K| tables, evaluated in FP64, cuviscoelas_3D.cu:72,188-190; the DC
// symbol of |k|^-1 is 0 because every use is followed by a derivative that
// vanishes there), inverse x-FFT, write Q1, Q3.
struct XPassArgs {
    int nx, ny, nz;
    long long sx, sy;
    int Pc, ztiles;
    long long shot_stride;
    FftPlan plan;
    const float2* tw;
    const float *kfx, *kfy, *kfz;
    double invN;
    int ncomp;
    const float2* P[3];
    float2* Q1[3];
    float2* Q3[3];
};

__global__ __launch_bounds__(kThreads) void k_xpass(XPassArgs a)
{
    __shared__ float2 tile[kTileMax / 2 + 64];
    __shared__ float2 work[kTileMax / 2 + 64];
    __shared__ float2 tws[kMaxN];
    const int shot = blockIdx.y;
    const int iy = blockIdx.x / a.ztiles;
    const int z0 = (blockIdx.x - iy * a.ztiles) * a.Pc;
    const int n = a.nx, Pc = a.Pc, ps = n + 1, T = n * Pc, nz = a.nz;
    const long long obase = (long long)iy * a.sy;
    const long long soff = (long long)shot * a.shot_stride;
    for (int t = threadIdx.x; t < n; t += kThreads) tws[t] = a.tw[t];
    const double ky = a.kfy[iy];
    for (int comp = 0; comp < a.ncomp; ++comp) {
        const float2* src = a.P[comp] + soff + obase;
        __syncthreads();
        for (int e = threadIdx.x; e < T; e += kThreads) {
            const int i = e / Pc, c = e - i * Pc, zc = z0 + c;
            tile[c * ps + i] = zc < nz ? src[i * a.sx + zc] : make_float2(0.f, 0.f);
        }
        __syncthreads();
        fft_tile(tile, tws, a.plan, Pc, ps, -1.f);
        for (int op = 0; op < 2; ++op) {
            for (int e = threadIdx.x; e < T; e += kThreads) {
                const int c = e / n, i = e - c * n;
                const int zc = z0 + c < nz ? z0 + c : nz - 1;
                const double kx = a.kfx[i], kz = a.kfz[zc];
                const double k2 = kx * kx + ky * ky + kz * kz;
                double s;
                if (op == 0) s = k2 > 0.0 ? a.invN / sqrt(k2) : 0.0;
                else s = a.invN * sqrt(k2);
                const float fs = (float)s;
                const float2 v = tile[c * ps + i];
                work[c * ps + i] = make_float2(v.x * fs, v.y * fs);
            }
            __syncthreads();
            fft_tile(work, tws, a.plan, Pc, ps, 1.f);
            float2* dst = (op == 0 ? a.Q1[comp] : a.Q3[comp]) + soff + obase;
            for (int e = threadIdx.x; e < T; e += kThreads) {
                const int i = e / Pc, c = e - i * Pc, zc = z0 + c;
                if (zc < nz) dst[i * a.sx + zc] = work[c * ps + i];
            }
            __syncthreads();
        }
    }
}

// Stress update (cuviscoelas_3D.cu:800-1015, generalised to VTI moduli) fused
// with the z-derivatives of the 15 fields, the stress C-PML memory variables
// on the face slabs, and the source injection (cuviscoelas_3D.cu:771-777).
constexpr int kSzCells = 4;  // cells per thread

struct StressZArgs {
    int nx, ny, nz, pml;
    int Pc;  // complex pencils per sub-tile (2*Pc rows)
    long long nrows;
    float dt;
    long long shot_stride, slabx_stride, slaby_stride, slabz_stride;
    FftPlan plan;
    const float2* tw;
    const float2 *symp, *symm;
    const float *pax, *pbx, *paxh, *pbxh;
    const float *pay, *pby, *payh, *pbyh;
    const float *paz, *pbz, *pazh, *pbzh;
    const float* fam[3][5];  // [component][A1 v, v, A3 v, A1 dv, dv]
    const float *ex[5], *gx[5], *hx[5];
    const float *ey[5], *gy[5], *hy[5];
    const float *c11[5], *c33[5], *c13[5], *cs[5], *csb[5];
    float* sig[6];  // pxx pyy pzz pxy pxz pyz
    float* phx[5];
    float* phy[5];
    float* phz[4];
    const int* src_kind;
    const long long* src_ip;
    const float* rik;
    const int* it;
};

__global__ __launch_bounds__(kThreads) void k_stress_z(StressZArgs a)
{
    __shared__ float2 tile[3 * (kTileMax / 8 + 32)];
    __shared__ float2 tws[kMaxN];
    const int shot = blockIdx.y;
    const int n = a.nz, Pc = a.Pc, ps = n + 1, T = 2 * Pc * n;
    const int sub = Pc * ps;  // float2 per sub-tile
    const long long r0 = (long long)blockIdx.x * 2 * Pc;
    const long long soff = (long long)shot * a.shot_stride;
    for (int t = threadIdx.x; t < n; t += kThreads) tws[t] = a.tw[t];
    const float* tf = reinterpret_cast<const float*>(tile);

    float acc[kSzCells][14];
#pragma unroll
    for (int q = 0; q < kSzCells; ++q)
#pragma unroll
        for (int k = 0; k < 14; ++k) acc[q][k] = 0.f;

    for (int j = 0; j < 5; ++j) {
        __syncthreads();
        load_row_pairs(tile, a.fam[2][j] + soff, r0, a.nrows, n, Pc, ps);
        load_row_pairs(tile + sub, a.fam[0][j] + soff, r0, a.nrows, n, Pc, ps);
        load_row_pairs(tile + 2 * sub, a.fam[1][j] + soff, r0, a.nrows, n, Pc, ps);
        __syncthreads();
        fft_tile(tile, tws, a.plan, 3 * Pc, ps, -1.f);
        for (int e = threadIdx.x; e < 3 * Pc * n; e += kThreads) {
            const int c = e / n, i = e - c * n;
            const float2 s = c < Pc ? a.symp[i] : a.symm[i];
            tile[c * ps + i] = cmul(tile[c * ps + i], s);
        }
        __syncthreads();
        fft_tile(tile, tws, a.plan, 3 * Pc, ps, 1.f);
#pragma unroll
        for (int q = 0; q < kSzCells; ++q) {
            const int e = threadIdx.x + q * kThreads;
            if (e >= T) continue;
            const int r = e / n, iz = e - r * n;
            const long long row = r0 + r;
            if (row >= a.nrows) continue;
            const long long ip = row * n + iz;
            const int lo = 2 * ((r >> 1) * ps + iz) + (r & 1);
            const float ez = tf[lo], gz = tf[2 * sub + lo], hz = tf[4 * sub + lo];
            const long long g = soff + ip;
            const float ex = a.ex[j][g], gx = a.gx[j][g], hx = a.hx[j][g];
            const float ey = a.ey[j][g], gy = a.gy[j][g], hy = a.hy[j][g];
            const float c11 = a.c11[j][ip], c33 = a.c33[j][ip], c13 = a.c13[j][ip];
            const float cs = a.cs[j][ip], csb = a.csb[j][ip];
            const float c12 = c11 - 2 * cs, c13d = c13 - 2 * cs;
            float* A = acc[q];
            A[0] += c11 * ex;   // s11x
            A[1] += c12 * ex;   // s12x
            A[2] += c13d * ex;  // s13x
            A[3] += c12 * ey;   // s12y
            A[4] += c11 * ey;   // s11y
            A[5] += c13d * ey;  // s13y
            A[6] += c13d * ez;  // s13z
            A[7] += c33 * ez;   // s33z
            A[8] += csb * gz;   // sb_xz  (dz vx)
            A[9] += csb * gx;   // sb_zx  (dx vz)
            A[10] += csb * hy;  // sb_xy  (dy vx)
            A[11] += csb * hx;  // sb_yx  (dx vy)
            A[12] += csb * hz;  // sb_yz  (dz vy)
            A[13] += csb * gy;  // sb_zy  (dy vz)
        }
    }

    const int pml = a.pml;
    const float dt = a.dt;
    const int it = *a.it;
    const int kind = a.src_kind[shot];
    const long long sip = a.src_ip[shot];
#pragma unroll
    for (int q = 0; q < kSzCells; ++q) {
        const int e = threadIdx.x + q * kThreads;
        if (e >= T) continue;
        const int r = e / n, iz = e - r * n;
        const long long row = r0 + r;
        if (row >= a.nrows) continue;
        const int ix = (int)(row / a.ny), iy = (int)(row - (long long)ix * a.ny);
        const long long ip = row * n + iz;
        const long long g = soff + ip;
        if (ix < a.nx - 1 && iy < a.ny - 1 && iz < n - 1) {
            const float* A = acc[q];
            float fxx_x = 0.f, fxy_x = 0.f, fxz_x = 0.f, fzx = 0.f, fyx = 0.f;
            float fyx_y = 0.f, fyy_y = 0.f, fyz_y = 0.f, fxy = 0.f, fzy = 0.f;
            float fzx_z = 0.f, fzz_z = 0.f, fxz = 0.f, fyz = 0.f;
            if (ix < pml || ix >= a.nx - pml) {
                const int si = ix < pml ? ix : ix - (a.nx - 2 * pml);
                const long long s = (long long)shot * a.slabx_stride + ((long long)si * a.ny + iy) * n + iz;
                const float Ax = a.pax[ix], Bx = a.pbx[ix], Ah = a.paxh[ix], Bh = a.pbxh[ix];
                float* p;
                p = a.phx[0] + s; fxx_x = *p + dt * (-Ax * *p - Bx * A[0]); *p = fxx_x;
                p = a.phx[1] + s; fxy_x = *p + dt * (-Ax * *p - Bx * A[1]); *p = fxy_x;
                p = a.phx[2] + s; fxz_x = *p + dt * (-Ax * *p - Bx * A[2]); *p = fxz_x;
                p = a.phx[3] + s; fzx = *p + dt * (-Ah * *p - Bh * A[9]); *p = fzx;
                p = a.phx[4] + s; fyx = *p + dt * (-Ah * *p - Bh * A[11]); *p = fyx;
            }
            if (iy < pml || iy >= a.ny - pml) {
                const int si = iy < pml ? iy : iy - (a.ny - 2 * pml);
                const long long s = (long long)shot * a.slaby_stride + ((long long)ix * 2 * pml + si) * n + iz;
                const float Ay = a.pay[iy], By = a.pby[iy], Ah = a.payh[iy], Bh = a.pbyh[iy];
                float* p;
                p = a.phy[0] + s; fyx_y = *p + dt * (-Ay * *p - By * A[3]); *p = fyx_y;
                p = a.phy[1] + s; fyy_y = *p + dt * (-Ay * *p - By * A[4]); *p = fyy_y;
                p = a.phy[2] + s; fyz_y = *p + dt * (-Ay * *p - By * A[5]); *p = fyz_y;
                p = a.phy[3] + s; fxy = *p + dt * (-Ah * *p - Bh * A[10]); *p = fxy;
                p = a.phy[4] + s; fzy = *p + dt * (-Ah * *p - Bh * A[13]); *p = fzy;
            }
            if (iz < pml || iz >= n - pml) {
                const int si = iz < pml ? iz : iz - (n - 2 * pml);
                const long long s = (long long)shot * a.slabz_stride + row * 2 * pml + si;
                const float Az = a.paz[iz], Bz = a.pbz[iz], Ah = a.pazh[iz], Bh = a.pbzh[iz];
                float* p;
                p = a.phz[0] + s; fzx_z = *p + dt * (-Az * *p - Bz * A[6]); *p = fzx_z;
                p = a.phz[1] + s; fzz_z = *p + dt * (-Az * *p - Bz * A[7]); *p = fzz_z;
                p = a.phz[2] + s; fxz = *p + dt * (-Ah * *p - Bh * A[8]); *p = fxz;
                p = a.phz[3] + s; fyz = *p + dt * (-Ah * *p - Bh * A[12]); *p = fyz;
            }
            float* pxx = a.sig[0] + g;
            float* pyy = a.sig[1] + g;
            float* pzz = a.sig[2] + g;
            float* pxy = a.sig[3] + g;
            float* pxz = a.sig[4] + g;
            float* pyz = a.sig[5] + g;
            *pxx = *pxx + dt * (A[0] + A[3] + A[6] + (fxx_x + fyx_y + fzx_z));
            *pyy = *pyy + dt * (A[1] + A[4] + A[6] + (fxy_x + fyy_y + fzx_z));
            *pzz = *pzz + dt * (A[2] + A[5] + A[7] + (fxz_x + fyz_y + fzz_z));
            *pxz = *pxz + dt * (A[8] + fxz + A[9] + fzx);
            *pyz = *pyz + dt * (A[12] + fyz + A[13] + fzy);
            *pxy = *pxy + dt * (A[10] + fxy + A[11] + fyx);
        }
        if (ip == sip) {
            // source after the stress update; the owner of the cell is the
            // only workgroup that reads this row of v (its z-derivative was
            // taken in the loop above), so the write cannot race.
            const float s = a.rik[it];
            if (kind == 3) {
                a.sig[0][g] += s;
                a.sig[1][g] += s;
                a.sig[2][g] += s;
            } else if (kind >= 0 && kind < 3) {
                float* vc = const_cast<float*>(a.fam[kind][1]);
                vc[g] += s;
            }
        }
    }
}

// Receivers: rec[it][r] = v[cell_r] for the three components (after the
// source injection), cuviscoelas_3D.cu:782-792.  Also advances nothing: the
// step counter is advanced by k_tick at the start of the next step.
struct RecArgs {
    const float *vx, *vy, *vz;
    long long shot_stride;
    const long long* rec_ip;  // all shots' receivers back to back
    const int* rec_off;       // per shot: first receiver, [nshots+1]
    float *rx, *ry, *rz;      // records, per shot offset rec_off[s]*nt
    int nt;
    const int* it;
};

__global__ __launch_bounds__(kThreads) void k_record(RecArgs a)
{
    const int shot = blockIdx.y;
    const int r0 = a.rec_off[shot], nr = a.rec_off[shot + 1] - r0;
    const int r = blockIdx.x * kThreads + threadIdx.x;
    if (r >= nr) return;
    const int it = *a.it;
    const long long ip = a.rec_ip[r0 + r] + (long long)shot * a.shot_stride;
    const long long o = (long long)r0 * a.nt + (long long)it * nr + r;
    a.rx[o] = a.vx[ip];
    a.ry[o] = a.vy[ip];
    a.rz[o] = a.vz[ip];
}

__global__ void k_tick(int* it) { *it = *it + 1; }

// Interior-window snapshot of one field when it % every == 0.
__global__ __launch_bounds__(kThreads) void k_snapshot(const float* f, int nx, int ny, int nz, int pml, int every,
                                                       const int* it, float* out, long long cap)
{
    const int t = *it;
    if (every <= 0 || t % every != 0) return;
    const long long k = t / every;
    const int wx = nx - 2 * pml, wy = ny - 2 * pml, wz = nz - 2 * pml;
    const long long W = (long long)wx * wy * wz;
    if ((k + 1) * W > cap) return;
    for (long long q = (long long)blockIdx.x * kThreads + threadIdx.x; q < W; q += (long long)gridDim.x * kThreads) {
        const int ix = (int)(q / ((long long)wy * wz));
        const long long rem = q - (long long)ix * wy * wz;
        const int iy = (int)(rem / wz), iz = (int)(rem - (long long)iy * wz);
        out[k * W + q] = f[((long long)(ix + pml) * ny + iy + pml) * nz + iz + pml];
    }
}

// Blow-up probe: max |v| over a field (bit pattern of a non-negative float
// orders like an int; NaN sorts above Inf).
__global__ __launch_bounds__(kThreads) void k_probe(const float* f, long long N, int* out)
{
    int m = 0;
    for (long long q = (long long)blockIdx.x * kThreads + threadIdx.x; q < N; q += (long long)gridDim.x * kThreads) {
        const int b = __float_as_int(fabsf(f[q]));
        m = b > m ? b : m;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const int o = __shfl_down(m, off);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// ---- setup kernels (once per model) ------------------------------------

// Q coefficients (cuviscoelas_3D.cu:195-205) in FP64, rounded to FP32.
// set: 0 P (vp, gp), 1 S (vs, gs), 2 H (vp*sqrt(1+2eps), gp), 3 D (v_delta, gp).
// The tau terms act on dv = v - v_prev, so they carry the 1/dt of the
// reference's (V - V_prev)/dt.
__device__ __forceinline__ void q_coeffs(double rho, double v, double g, double w0, double dt, double* o)
{
    const double pi = 3.1415926;
    const double c = v * cos(g * pi / 2.0);
    o[0] = -rho * c * g * w0;
    o[1] = rho * (c * c);
    o[2] = rho * pow(c, 3.0) * g / w0;
    o[3] = rho * c * g * pi / dt;
    o[4] = rho * (c * c) * (g * g) * pi / w0 / dt;
}

struct CoefArgs {
    long long N;
    int vti;
    double w0, dt;
    const float *vp, *gp, *vs, *gs, *rho, *eps, *del;
    float* out;  // [nsets*5][N]: P, S, (H, D)
};

__global__ __launch_bounds__(kThreads) void k_coefficients(CoefArgs a)
{
    for (long long ip = (long long)blockIdx.x * kThreads + threadIdx.x; ip < a.N; ip += (long long)gridDim.x * kThreads) {
        double o[5];
        const double r = a.rho[ip];
        q_coeffs(r, a.vp[ip], a.gp[ip], a.w0, a.dt, o);
        for (int j = 0; j < 5; ++j) a.out[(0 * 5 + j) * a.N + ip] = (float)o[j];
        q_coeffs(r, a.vs[ip], a.gs[ip], a.w0, a.dt, o);
        for (int j = 0; j < 5; ++j) a.out[(1 * 5 + j) * a.N + ip] = (float)o[j];
        if (a.vti) {
            const double e = a.eps[ip], dl = a.del[ip], vp = a.vp[ip], vs = a.vs[ip];
            const double vh = e == 0.0 ? vp : vp * sqrt(1.0 + 2.0 * e);
            q_coeffs(r, vh, a.gp[ip], a.w0, a.dt, o);
            for (int j = 0; j < 5; ++j) a.out[(2 * 5 + j) * a.N + ip] = (float)o[j];
            double vd = vp;
            if (dl != 0.0) {
                const double c33 = r * vp * vp, c44 = r * vs * vs;
                const double prod = (c33 - c44) * (c33 * (1.0 + 2.0 * dl) - c44);
                vd = sqrt((sqrt(prod > 0.0 ? prod : 0.0) + c44) / r);
            }
            q_coeffs(r, vd, a.gp[ip], a.w0, a.dt, o);
            for (int j = 0; j < 5; ++j) a.out[(3 * 5 + j) * a.N + ip] = (float)o[j];
        }
    }
}

// 8-corner cube average of the S set (cuviscoelas_3D.cu:903-907) and the
// buoyancies dt / (0.5 (rho + rho_next)) (cuviscoelas_3D.cu:1077-1092).
struct AvgArgs {
    int nx, ny, nz;
    long long sx, sy, N;
    float dt;
    const float* S;  // [5][N]
    float* Sb;       // [5][N]
    const float* rho;
    float* buoy;     // [3][N]
};

__global__ __launch_bounds__(kThreads) void k_averages(AvgArgs a)
{
    for (long long ip = (long long)blockIdx.x * kThreads + threadIdx.x; ip < a.N; ip += (long long)gridDim.x * kThreads) {
        const int ix = (int)(ip / a.sx);
        const int iy = (int)((ip - (long long)ix * a.sx) / a.sy);
        const int iz = (int)(ip - (long long)ix * a.sx - (long long)iy * a.sy);
        const bool in = ix < a.nx - 1 && iy < a.ny - 1 && iz < a.nz - 1;
        for (int j = 0; j < 5; ++j) {
            const float* c = a.S + (long long)j * a.N;
            float v = 0.f;
            if (in) {
                const float s8 = c[ip] + c[ip + 1] + c[ip + a.sy] + c[ip + 1 + a.sy] + c[ip + a.sx] + c[ip + a.sx + 1]
                               + c[ip + a.sx + a.sy] + c[ip + a.sx + 1 + a.sy];
                v = 0.125f * s8;
            }
            a.Sb[(long long)j * a.N + ip] = v;
        }
        // rho + rho' is an FP32 sum in the reference, the rest FP64
        const float r = a.rho[ip];
        a.buoy[ip] = ix < a.nx - 1 ? (float)((double)a.dt / (0.5 * (double)(r + a.rho[ip + a.sx]))) : 0.f;
        a.buoy[a.N + ip] = iy < a.ny - 1 ? (float)((double)a.dt / (0.5 * (double)(r + a.rho[ip + a.sy]))) : 0.f;
        a.buoy[2 * a.N + ip] = iz < a.nz - 1 ? (float)((double)a.dt / (0.5 * (double)(r + a.rho[ip + 1]))) : 0.f;
    }
}

}  // namespace qvti
