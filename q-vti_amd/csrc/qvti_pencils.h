// qvti_pencils.h -- pencil kernels of the time step on the two-level register
// FFT (qvti_fft2l.h).  Every kernel is a template over the split G = L2<N1,N2,PW>
// of its axis length; the row (z) kernels run 256 threads, the strided-axis
// (x, y) kernels 256*PW (the tile width the autotuner picks).
//
// Per step (reference loop: cuviscoelas_3D.cu:1255-1605):
//   velocity half-step   k2_deriv_strided(x, VEL_FIRST; advances the step counter)
//                        -> k2_deriv_strided(y, VEL_ACC)
//                        -> k2_vel_z (z-derivatives, velocity update, forward z-FFT of v + i dv)
//   stress half-step     k2_cfft_strided(y, fwd) -> k2_xpass -> k2_cfft_strided(y, inv)
//                        -> k2_zrows (inverse z) -> k2_deriv_strided(x, STORE) -> k2_deriv_strided(y, STORE)
//                        -> k2_stress_z (z-derivatives, stress update, source injection)
//   receivers            k_record
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
//    only A1 (v + i dv) and A3 v need the full 3-D transform.  v and dv are
//    packed together because their magnitudes are comparable (|dv| ~ w dt |v|);
//    packing v with dv/dt instead lets the FP32 rounding of the larger part
//    swamp the smaller one.
#pragma once
#include "qvti_fft2l.h"
#include "qvti_layout.h"

namespace qvti {

constexpr int kThreads = 256;
#ifndef QVTI_STRIDED_WAVES
#define QVTI_STRIDED_WAVES 1
#endif
#ifndef QVTI_XPASS_WAVES
#define QVTI_XPASS_WAVES 1
#endif
// minimum waves per SIMD the z-row kernels are compiled for (1: no bound on
// their registers; 4 caps them at 128 VGPRs, the occupancy their LDS allows)
#ifndef QVTI_VELZ_WAVES
#define QVTI_VELZ_WAVES 1
#endif
#ifndef QVTI_STRESSZ_WAVES
#define QVTI_STRESSZ_WAVES 1
#endif
#ifndef QVTI_VZ_CHUNK
#define QVTI_VZ_CHUNK 4
#endif
#ifndef QVTI_VEL_CHUNK
#define QVTI_VEL_CHUNK 4
#endif
enum { MODE_STORE = 0, MODE_VEL_FIRST = 1, MODE_VEL_ACC = 2 };

// XCD-aware placement of (row tile, shot) on a 1-D grid of
// ceil(ntiles / (8 run)) * 8 run * nshots workgroups.  Workgroup ids are dealt
// round-robin to the 8 XCDs; id = ((c run + k) nshots + shot) 8 + x runs tile
// (8 c + x) run + k on XCD x.  So
//  * the shots of one row tile run on the same XCD (same L2), dispatched
//    close together, and the model fields they all read (coefficients,
//    buoyancies) come from cache after the first shot;
//  * `run` consecutive row tiles run on one XCD, one after another, so the
//    128-byte lines (and 64-cell blocks) that neighbouring tiles share when a
//    tile does not end on a line or block boundary are fetched from HBM once.
// Returns tile = -1 for the padding workgroups.
struct TileShot {
    int tile, shot;
};
__device__ __forceinline__ TileShot tile_shot(int ntiles, int nshots, int run)
{
    const int id = blockIdx.x;
    const int x = id & 7, q = id >> 3;
    const int ck = q / nshots, shot = q - ck * nshots;
    const int c = ck / run, k = ck - c * run;
    const int t = (c * 8 + x) * run + k;
    return {t < ntiles ? t : -1, shot};
}

// (plane, z-tile) of a strided-axis workgroup.  run = 0: blockIdx.x =
// plane * ztiles + tile.  run >= 1: XCD-aware placement on a 1-D grid of
// ceil(nplanes / (8 run)) * 8 run * ztiles workgroups, id = ((c run + k)
// ztiles + zt) 8 + x running (plane (8 c + x) run + k, z-tile zt) on XCD x:
// the z-tiles of one plane, and `run` consecutive planes, run on the same XCD
// (same L2) one after another, so the 128-byte lines that neighbouring tiles
// share - within a row when a tile's rows do not end on a line boundary, and
// between the end of a row and the start of the next plane's row - are
// fetched from HBM once.  plane = -1 for the padding workgroups.
struct PlaneTile {
    int plane, zt;
};
__device__ __forceinline__ PlaneTile plane_tile(int ztiles, int nplanes, int run)
{
    const int id = blockIdx.x;
    if (!run) {
        const int pl = id / ztiles;
        return {pl, id - pl * ztiles};
    }
    const int x = id & 7, q = id >> 3;
    const int pk = q / ztiles, zt = q - pk * ztiles;
    const int c = pk / run, k = pk - c * run;
    const int pl = (c * 8 + x) * run + k;
    return {pl < nplanes ? pl : -1, zt};
}

template <class G>
__device__ __forceinline__ void load_twiddles(float2* tws, const float2* tw)
{
    for (int t = threadIdx.x; t < G::TWN; t += G::T) twiddle_entry<G>(tws, tw, t);
}

// ------------------------------------------------------------------------
// 1-D shift-derivatives of real fields along a strided axis (x or y).  Tile:
// G::P complex pencils = 2*G::P adjacent real z-columns at one value of the
// other in-plane index, loaded as float2 (two real pencils per complex one).
// Modes: STORE out = d; VEL_FIRST out = d + phi; VEL_ACC out = acc + d + phi,
// with the velocity C-PML memory variable phi updated on the axis' face slabs
// (cuviscoelas_3D.cu:1049-1074).
struct DJob {
    const float* in;
    float* out;
    const float* acc;
    float* phi;
    int sym;   // 0: D+, 1: D-
    int half;  // 1: half-cell averaged C-PML a, b
};

struct StridedArgs {
    int nx, ny, nz, pml;
    long long sx, sy;
    int axis;  // 0: x (stride sx), 1: y (stride sy)
    int ztiles, nplanes, run;  // workgroup placement: plane_tile()
    int mode;
    float dt;
    long long shot_stride, slab_stride;
    const float2* tw;
    const float2* sym[2];
    const float *pa, *pb, *pah, *pbh;
    int njobs;
    // STORE mode output layout: 0 = natural [shot][N] per job; nf > 0 =
    // blocked group of nf fields per shot ([shot][N/64][nf][64], see
    // blk_index), each job's out pointing at its field (group base + f*64)
    int bnf;
    // per-shot length of a blocked group's field: N rounded up to whole
    // 64-cell blocks (shot s of an nf-field group starts at s * nf * blk_stride)
    long long blk_stride;
    // step counter advanced by this launch (the first kernel of a step), or null
    int* tick;
    DJob job[15];
};

template <class G, int MODE>
__global__ __launch_bounds__(G::T) __attribute__((amdgpu_waves_per_eu(QVTI_STRIDED_WAVES)))
void k2_deriv_strided(StridedArgs a)
{
    __shared__ float2 lds[G::P * G::PS];
    __shared__ float2 tws[G::TWN];
    const int tid = threadIdx.x, shot = blockIdx.y;
    // one thread of the grid advances the step counter; the kernels of the
    // step that read it run after this one
    if (a.tick != nullptr && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && tid == 0) *a.tick = *a.tick + 1;
    const PlaneTile pt = plane_tile(a.ztiles, a.nplanes, a.run);
    if (pt.plane < 0) return;
    const int other = pt.plane;
    const int p = tid % G::P, jj = tid / G::P;
    const int zc = pt.zt * 2 * G::P + 2 * p;
    const int nz = a.nz;
    const bool zok = zc < nz;
    // columns past nz (last tile of a ragged z) load a valid column and are
    // never stored, so the loads need no branch
    const int zl = zok ? zc : nz - 2;
    const bool act1 = jj < G::N2, act2 = jj < G::N1;
    // uniform per-block base (SGPRs) + 32-bit per-thread element offsets
    const int astride = (int)(a.axis == 0 ? a.sx : a.sy);
    const long long poff = a.axis == 0 ? (long long)other * a.sy : (long long)other * a.sx;  // plane offset in the shot
    const long long ubase = (long long)shot * a.shot_stride + poff;
    float2* pen = lds + p * G::PS;
    load_twiddles<G>(tws, a.tw);
    for (int jb = blockIdx.z; jb < a.njobs; jb += gridDim.z) {
        const DJob J = a.job[jb];
        __syncthreads();
        if (act1) {
            float2 v[G::N1];
            // uniform row pointer (SGPRs) + one unsigned 32-bit lane offset
            const unsigned voff = (unsigned)(jj * astride + zl);
#pragma unroll
            for (int j1 = 0; j1 < G::N1; ++j1) {
                const float* __restrict__ rp = J.in + ubase + (long long)(j1 * G::N2) * astride;
                v[j1] = *reinterpret_cast<const float2*>(rp + voff);
            }
            fwd_l1<G>(v, pen, tws, jj);
        }
        __syncthreads();
        float2 b[G::N2];
        if (act2) {
            fwd_l2<G>(b, pen, jj);
            const float2* __restrict__ S = a.sym[J.sym];
#pragma unroll
            for (int k2 = 0; k2 < G::N2; ++k2) b[k2] = c_mul(b[k2], S[jj + G::N1 * k2]);
        }
        __syncthreads();
        if (act2) inv_la<G>(b, pen, tws + G::T2OFF, jj);
        __syncthreads();
        const unsigned voff = (unsigned)(jj * astride + zc);
        if (MODE == MODE_STORE) {
            if (act1 && zok) {
                float2 d[G::N1];
                inv_lb<G>(d, pen, jj);
                if (a.bnf) {
                    // blocked group: (ip, ip+1) share a 64-cell block (ip even)
                    float* ob = J.out + (long long)shot * a.bnf * a.blk_stride;
                    const unsigned ip0 = (unsigned)(poff + jj * astride + zc);
                    const unsigned bs = (unsigned)(kBlk * a.bnf);
#pragma unroll
                    for (int j1 = 0; j1 < G::N1; ++j1) {
                        const unsigned ip = ip0 + (unsigned)(G::N2 * j1 * astride);
                        *reinterpret_cast<float2*>(ob + ((ip >> 6) * bs + (ip & 63u))) = d[j1];
                    }
                } else {
#pragma unroll
                    for (int j1 = 0; j1 < G::N1; ++j1) {
                        const long long urow = ubase + (long long)(G::N2 * j1) * astride;
                        *reinterpret_cast<float2*>(J.out + urow + voff) = d[j1];
                    }
                }
            }
        } else {
            // The C-PML branch differs per element: park the pencil's result in
            // its own LDS slots (natural order) and finish element by element,
            // which keeps the register footprint at the FFT's.
            float2 d[G::N1];
            if (act1 && zok) inv_lb<G>(d, pen, jj);
            __syncthreads();
            if (act1 && zok) {
#pragma unroll
                for (int j1 = 0; j1 < G::N1; ++j1) pen[jj + G::N2 * j1] = d[j1];
                const int pml = a.pml, n = G::N;
                // an opaque copy of the lane's row index: keeps the per-element
                // C-PML indices and table loads inside the job loop instead of
                // being hoisted (and held in registers through the FFTs)
                int jl = jj;
                asm volatile("" : "+v"(jl));
                const float* __restrict__ pa = J.half ? a.pah : a.pa;
                const float* __restrict__ pb = J.half ? a.pbh : a.pb;
                float2* __restrict__ phs = reinterpret_cast<float2*>(J.phi + (long long)shot * a.slab_stride);
                // Chunks of CH elements: every load of a chunk (unconditional
                // and in bounds - a non-slab element reads slab plane 0) before
                // its stores, so the chunk's memory round trips overlap.
                constexpr int CH = G::N1 % QVTI_VEL_CHUNK == 0 ? QVTI_VEL_CHUNK : 1;
#pragma unroll
                for (int c0 = 0; c0 < G::N1; c0 += CH) {
                    float2 rr[CH], q[CH];
                    int sidx[CH];
                    bool upd[CH];
#pragma unroll
                    for (int jc = 0; jc < CH; ++jc) {
                        const int j1 = c0 + jc;
                        const int i = jl + G::N2 * j1;
                        const long long urow = ubase + (long long)(G::N2 * j1) * astride;
                        float2 r = pen[i];
                        const int ix = a.axis == 0 ? i : other;
                        const int iy = a.axis == 0 ? other : i;
                        upd[jc] = (i < pml || i >= n - pml) && ix < a.nx - 1 && iy < a.ny - 1;
                        const int si = i < pml ? i : (i >= n - pml ? i - (n - 2 * pml) : 0);
                        sidx[jc] = (a.axis == 0 ? (si * a.ny + iy) * nz + zc : (ix * 2 * pml + si) * nz + zc) / 2;
                        const float2 q0 = phs[sidx[jc]];
                        const float A = pa[i], B = pb[i];
                        q[jc] = q0;
                        if (upd[jc]) {
                            if (zc < nz - 1) { q[jc].x = q0.x + a.dt * (-A * q0.x - B * r.x); r.x = r.x + q[jc].x; }
                            if (zc + 1 < nz - 1) { q[jc].y = q0.y + a.dt * (-A * q0.y - B * r.y); r.y = r.y + q[jc].y; }
                        }
                        if (MODE == MODE_VEL_ACC) {
                            const float2 ac = *reinterpret_cast<const float2*>(J.acc + urow + voff);
                            r = make_float2(ac.x + r.x, ac.y + r.y);
                        }
                        rr[jc] = r;
                    }
#pragma unroll
                    for (int jc = 0; jc < CH; ++jc) {
                        const int j1 = c0 + jc;
                        const long long urow = ubase + (long long)(G::N2 * j1) * astride;
                        if (upd[jc]) phs[sidx[jc]] = q[jc];
                        *reinterpret_cast<float2*>(J.out + urow + voff) = rr[jc];
                    }
                    asm volatile("" ::: "memory");  // the next chunk's loads stay after this chunk
                }
            }
        }
    }
}

// ------------------------------------------------------------------------
// Velocity update (cuviscoelas_3D.cu:1018-1101) fused with the z-derivatives
// and with the first (z) pass of the stress step's 3-D transform:
//   v += b * (acc_xy + dz sigma + phi_z);  dv = v_new - v_old + src_prev;
//   P0 = FFT_z(vx + i vy), P1 = FFT_z(dvx + i dvy), P2 = FFT_z(vz + i dvz)
// dv is the change since the previous step's pre-injection velocity (the
// reference's V - V_prev, cuviscoelas_3D.cu:517-522, 1124-1131, without 1/dt).
// Tile: G::P rows (z-pencils).
struct VelZArgs {
    int nx, ny, nz, pml;
    int ntiles, nshots, run;  // grid: tile_shot()
    long long nrows;
    float dt;
    long long shot_stride, slab_stride;
    const float2* tw;
    const float2 *symp, *symm;
    const float *pa, *pb, *pah, *pbh;
    const float *pxz, *pyz, *pzz;
    const float* acc[3];
    float* v[3];
    float* dv[3];
    float2* P[3];
    const float* buoy[3];
    float* phi[3];  // z-slab: phi_pxz_z (node), phi_pyz_z (node), phi_pzz_z (half)
    const int* src_kind;
    const long long* src_ip;
    const float* rik;
    const int* it;
    // extra point sources added to v after the previous step (the receiver
    // side of a reverse-time migration): dense [3][shot][N], zero away from
    // the injection cells; they enter dv like the shot's own source
    // (k2_vel_z<G, true> only)
    const float* inj[3];
};

// Velocity update of one component on the thread's level-1 elements of one
// row (iz = j + N2*j1), from its z-derivative d (real or imaginary part):
// C-PML memory variable on the z-slab, v += b * (acc + dz + phi), dv = v_new
// - v_old (+ the previous step's source at the injection cell).  Writes v and
// dv (rows with rowdom / ok1 only), returns (v_new, dv).  The pointers are
// row bases (v, dv, acc, bu: the row; ph: the row's z-slab of 2*pml values)
// and never alias, which lets the compiler issue the loads of later elements
// before the stores of earlier ones.
template <class G, bool INJ>
__device__ __forceinline__ void vz_component(float2 (&out)[G::N1], const float2 (&d)[G::N1], const bool imag,
                                             float* __restrict__ v, float* __restrict__ dv,
                                             const float* __restrict__ acc, const float* __restrict__ bu,
                                             float* __restrict__ ph, const float* __restrict__ pa,
                                             const float* __restrict__ pb, const int j, const int pml,
                                             const float dt, const bool rowdom, const bool ok1, const int src_iz,
                                             const float sprev, const float* __restrict__ inj)
{
    constexpr int n = G::N;
    int jl = j;  // opaque per call: nothing is shared (and held) across components
    asm volatile("" : "+v"(jl));
    // all loads and arithmetic first (every load unconditional and in bounds:
    // ph is a valid slab row even where it is not used), all stores after, so
    // no load waits behind a store or a branch
    // (in chunks of QVTI_VZ_CHUNK elements, to bound the registers in flight)
    constexpr int CH = G::N1 % QVTI_VZ_CHUNK == 0 ? QVTI_VZ_CHUNK : G::N1;
#pragma unroll
    for (int c0 = 0; c0 < G::N1; c0 += CH) {
    float qn[CH];
#pragma unroll
    for (int jc = 0; jc < CH; ++jc) {
        const int j1 = c0 + jc;
        const int iz = jl + G::N2 * j1;
        const float vold = v[iz], bb = bu[iz], ac = acc[iz];
        const bool slab = iz < pml || iz >= n - pml;
        const int si = iz < pml ? iz : (iz >= n - pml ? iz - (n - 2 * pml) : 0);
        const float q0 = ph[si], A = pa[iz], B = pb[iz];
        const bool upd = rowdom && iz < n - 1;
        float dz = imag ? d[j1].y : d[j1].x;
        qn[jc] = q0 + dt * (-A * q0 - B * dz);
        if (upd && slab) dz = dz + qn[jc];
        const float vnew = upd ? vold + bb * (ac + dz) : vold;
        float dvv = vnew - vold;
        if (iz == src_iz) dvv += sprev;
        if constexpr (INJ) dvv += inj[iz];
        out[j1] = make_float2(vnew, dvv);
    }
#pragma unroll
    for (int jc = 0; jc < CH; ++jc) {
        const int j1 = c0 + jc;
        const int iz = jl + G::N2 * j1;
        const bool upd = rowdom && iz < n - 1;
        if (upd && (iz < pml || iz >= n - pml)) ph[iz < pml ? iz : iz - (n - 2 * pml)] = qn[jc];
        if (upd) v[iz] = out[j1].x;
        if (ok1) dv[iz] = out[j1].y;
    }
    asm volatile("" ::: "memory");  // the next chunk's loads stay after this chunk
    }
}

template <class G, bool INJ = false>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(QVTI_VELZ_WAVES))) void k2_vel_z(VelZArgs a)
{
    __shared__ float2 lds[G::P * G::PS];
    __shared__ float2 tws[G::TWN];
    constexpr int n = G::N;
    const TileShot ts = tile_shot(a.ntiles, a.nshots, a.run);
    if (ts.tile < 0) return;  // padding workgroup (whole workgroup, before any barrier)
    const int tid = threadIdx.x, shot = ts.shot;
    // level 1/B mapping (column j fastest) and level 2/A mapping (row k1 fastest)
    const int j = tid % G::N2, p1 = tid / G::N2;
    const int k1 = tid % G::N1, p2 = tid / G::N1;
    const bool act1 = p1 < G::P, act2 = p2 < G::P;
    const long long r0 = (long long)ts.tile * G::P;
    const long long r1 = r0 + p1, r2 = r0 + p2;
    const bool ok1 = act1 && r1 < a.nrows, ok2 = act2 && r2 < a.nrows;
    // uniform block base (shot + first row) and unsigned 32-bit lane offsets;
    // rows past the end (ragged last block) read row r0 and store nothing
    const long long ubase = (long long)shot * a.shot_stride + r0 * n;
    const unsigned lrow1 = ok1 ? (unsigned)(p1 * n) : 0u;
    float2* pen1 = lds + p1 * G::PS;
    float2* pen2 = lds + p2 * G::PS;
    load_twiddles<G>(tws, a.tw);

    // derivative of (fa + i fb) along z with symbol S, result in d (level-1 layout)
    auto deriv = [&](const float* fa, const float* fb, const float2* S, float2 (&d)[G::N1]) {
        __syncthreads();
        if (act1) {
            float2 v[G::N1];
#pragma unroll
            for (int j1 = 0; j1 < G::N1; ++j1) {
                const unsigned o = lrow1 + j1 * G::N2 + j;
                v[j1] = make_float2(fa[ubase + o], fb ? fb[ubase + o] : 0.f);
            }
            fwd_l1<G>(v, pen1, tws, j);
        }
        __syncthreads();
        float2 b[G::N2];
        if (act2) {
            fwd_l2<G>(b, pen2, k1);
#pragma unroll
            for (int k2 = 0; k2 < G::N2; ++k2) b[k2] = c_mul(b[k2], S[k1 + G::N1 * k2]);
        }
        __syncthreads();
        if (act2) inv_la<G>(b, pen2, tws + G::T2OFF, k1);
        __syncthreads();
        if (act1) inv_lb<G>(d, pen1, j);
    };

    const int it = *a.it;
    const int kind = a.src_kind[shot];
    const long long sip = a.src_ip[shot];
    const float sprev = it > 0 ? a.rik[it - 1] : 0.f;
    const int ix = (int)(r1 / a.ny), iy = (int)(r1 - (long long)ix * a.ny);
    const bool rowdom = ok1 && ix < a.nx - 1 && iy < a.ny - 1;
    const long long slab_row = (long long)shot * a.slab_stride + r1 * 2 * a.pml;

    // component c on this thread's elements of its row
    auto update = [&](float2 (&out)[G::N1], const float2 (&d)[G::N1], int c, bool imag) {
        const long long rb = ubase + lrow1;  // row base of this thread's row in the shot's fields
        const int src_iz = (kind == c && sip >= r1 * n && sip < (r1 + 1) * n) ? (int)(sip - r1 * n) : -1;
        const bool h = c == 2;
        vz_component<G, INJ>(out, d, imag, a.v[c] + rb, a.dv[c] + rb, a.acc[c] + rb,
                        a.buoy[c] + (rb - (long long)shot * a.shot_stride),
                        a.phi[c] + (ok1 ? slab_row : (long long)shot * a.slab_stride),
                        h ? a.pah : a.pa, h ? a.pbh : a.pb, j, a.pml, a.dt, rowdom, ok1, src_iz, sprev,
                        INJ ? a.inj[c] + rb : nullptr);
    };
    // forward z-FFT of x (first pass of the 3-D transform) into dst
    auto fft_store = [&](float2 (&x)[G::N1], float2* dstp) {
        __syncthreads();
        if (act1) fwd_l1<G>(x, pen1, tws, j);
        __syncthreads();
        if (ok2) {
            float2 b[G::N2];
            fwd_l2<G>(b, pen2, k1);
            float2* dst = dstp + ubase + (unsigned)(p2 * n);
#pragma unroll
            for (int k2 = 0; k2 < G::N2; ++k2) dst[k1 + G::N1 * k2] = b[k2];
        }
    };

    // Packing of the stress step's transforms: P0 = F(vx + i vy),
    // P1 = F(dvx + i dvy), P2 = F(vz + i dvz).  The |k| operator is then
    // needed on P0 and P2 only (A3 acts on v, not on dv).
    {
        float2 d[G::N1];
        deriv(a.pxz, a.pyz, a.symp, d);  // dz pxz (vx) + i dz pyz (vy), D+
        float2 u[G::N1];
        update(u, d, 0, false);
        // park (vx, dvx) in this thread's own pencil slots while vy is updated
        __syncthreads();  // every thread has read the pencil buffer (inv_lb)
        if (act1) {  // (threads past the last pencil own no slots)
#pragma unroll
            for (int j1 = 0; j1 < G::N1; ++j1) pen1[j + G::N2 * j1] = u[j1];
        }
        update(u, d, 1, true);
        float2 x0[G::N1], x1[G::N1];
#pragma unroll
        for (int j1 = 0; j1 < G::N1; ++j1) {
            const float2 ux = act1 ? pen1[j + G::N2 * j1] : make_float2(0.f, 0.f);
            x0[j1] = make_float2(ux.x, u[j1].x);
            x1[j1] = make_float2(ux.y, u[j1].y);
        }
        fft_store(x0, a.P[0]);
        fft_store(x1, a.P[1]);
    }
    {
        float2 d[G::N1];
        deriv(a.pzz, nullptr, a.symm, d);  // dz pzz (vz), D-
        float2 x2[G::N1];
        update(x2, d, 2, false);
        fft_store(x2, a.P[2]);
    }
}

// ------------------------------------------------------------------------
// In-place complex transforms along a strided axis (y: stride sy, x: stride
// sx) for up to 6 complex fields.  Tile: G::P adjacent complex z-columns.
// DIR -1: forward (natural-order input, natural-order spectrum out);
// DIR +1: inverse.
struct CStridedArgs {
    int nz;
    long long sx, sy;
    int axis, ztiles, nplanes, run;  // workgroup placement: plane_tile()
    long long shot_stride;
    const float2* tw;
    int nfields;
    float2* f[8];
};

template <class G, int DIR>
__global__ __launch_bounds__(G::T) void k2_cfft_strided(CStridedArgs a)
{
    __shared__ float2 lds[G::P * G::PS];
    __shared__ float2 tws[G::TWN];
    const int tid = threadIdx.x, shot = blockIdx.y;
    const PlaneTile pt = plane_tile(a.ztiles, a.nplanes, a.run);
    if (pt.plane < 0) return;
    const int other = pt.plane;
    const int p = tid % G::P, jj = tid / G::P;
    const int zc = pt.zt * G::P + p;
    const bool zok = zc < a.nz;
    const bool act1 = jj < G::N2 && zok, act2 = jj < G::N1 && zok;
    const int astride = (int)(a.axis == 0 ? a.sx : a.sy);
    // uniform tile base (shot, plane, first column) + unsigned lane offset
    const long long tbase = (long long)shot * a.shot_stride +
                            (a.axis == 0 ? (long long)other * a.sy : (long long)other * a.sx) +
                            (long long)pt.zt * G::P;
    const unsigned voff = (unsigned)(jj * astride + p);
    float2* pen = lds + p * G::PS;
    load_twiddles<G>(tws, a.tw);
    for (int f = blockIdx.z; f < a.nfields; f += gridDim.z) {
        float2* const fb = a.f[f] + tbase;
        __syncthreads();
        if (DIR < 0) {
            if (act1) {
                float2 v[G::N1];
#pragma unroll
                for (int j1 = 0; j1 < G::N1; ++j1) v[j1] = (fb + (long long)(j1 * G::N2) * astride)[voff];
                fwd_l1<G>(v, pen, tws, jj);
            }
            __syncthreads();
            if (act2) {
                float2 b[G::N2];
                fwd_l2<G>(b, pen, jj);
#pragma unroll
                for (int k2 = 0; k2 < G::N2; ++k2) (fb + (long long)(G::N1 * k2) * astride)[voff] = b[k2];
            }
        } else {
            if (act2) {
                float2 b[G::N2];
#pragma unroll
                for (int k2 = 0; k2 < G::N2; ++k2) b[k2] = (fb + (long long)(G::N1 * k2) * astride)[voff];
                inv_la<G>(b, pen, tws + G::T2OFF, jj);
            }
            __syncthreads();
            if (act1) {
                float2 d[G::N1];
                inv_lb<G>(d, pen, jj);
#pragma unroll
                for (int j1 = 0; j1 < G::N1; ++j1) (fb + (long long)(G::N2 * j1) * astride)[voff] = d[j1];
            }
        }
    }
}

// ------------------------------------------------------------------------
// x-pass of the 3-D transform: forward x-FFT of P (already z- and
// y-transformed), multiply by |k|^-1 / N and |k| / N (|k|^2 = kx^2+ky^2+kz^2
// from the folded 1-D |K| tables, evaluated in FP64 and rounded once,
// cuviscoelas_3D.cu:72, 188-190; the DC symbol of |k|^-1 is 0 because every
// use is followed by a derivative that vanishes there), inverse x-FFT,
// write Q1, Q3.
//
// KS (opt-in k-space correction and wavenumber filter): every stress
// operator A_j, A2 = 1 included, is multiplied by
//   F(|k|) = sinc^2(vh |k|) * W(|k|),  sinc(x) = sin(x)/x,  vh = vref dt / 2
// (the disabled sinc2nd of cuviscoelas_3D.cu:183-190), with W the radial
// Tukey taper: 1 for |k| <= k0, (1 + cos(pi (|k| - k0) / (kc - k0))) / 2 up
// to kc, 0 beyond (the filter the reference declares but never builds,
// Myfunctions.h:69, viscoelas_3D.cpp:111-112).  A2 F is no longer a 1-D
// operator, so a third product F P / N is inverse-transformed into Q2.
struct XPassArgs {
    int ny, nz;
    long long sx, sy;
    int ztiles, run;  // planes: ny; workgroup placement: plane_tile()
    long long shot_stride;
    const float2* tw;
    const float *kfx, *kfy, *kfz;
    double invN;
    int ncomp;
    const float2* P[3];
    float2* Q1[3];
    float2* Q3[3];
    float2* Q2[3];           // KS only
    float ks_vh;             // vref dt / 2 (0: no sinc^2 factor)
    float ks_k0, ks_kc;      // taper start and cut-off (kc = 0: no filter)
    float ks_pw;             // pi / (kc - k0) (0 when k0 == kc)
};

// F(|k|) of the KS mode, in FP32.
__device__ __forceinline__ float kspace_factor(const XPassArgs& a, float k)
{
    float f = 1.f;
    if (a.ks_vh > 0.f) {
        const float x = a.ks_vh * k;
        const float sx = x > 0.f ? sinf(x) / x : 1.f;
        f = sx * sx;
    }
    if (a.ks_kc > 0.f) {
        if (k > a.ks_kc) f = 0.f;
        else if (k > a.ks_k0) f *= 0.5f * (1.f + cosf(a.ks_pw * (k - a.ks_k0)));
    }
    return f;
}

template <class G, bool KS = false>
__global__ __launch_bounds__(G::T) __attribute__((amdgpu_waves_per_eu(QVTI_XPASS_WAVES))) void k2_xpass(XPassArgs a)
{
    __shared__ float2 lds[G::P * G::PS];
    __shared__ float2 tws[G::TWN];
    const int tid = threadIdx.x, shot = blockIdx.y;
    const PlaneTile pt = plane_tile(a.ztiles, a.ny, a.run);
    if (pt.plane < 0) return;
    const int iy = pt.plane;
    const int p = tid % G::P, jj = tid / G::P;
    const int zc = pt.zt * G::P + p;
    const bool zok = zc < a.nz;
    const bool act1 = jj < G::N2 && zok, act2 = jj < G::N1 && zok;
    // uniform tile base (shot, plane, first column) + unsigned lane offset
    const long long tbase = (long long)shot * a.shot_stride + (long long)iy * a.sy +
                            (long long)pt.zt * G::P;
    const int sxi = (int)a.sx;
    const unsigned voff = (unsigned)(jj * sxi + p);
    float2* pen = lds + p * G::PS;
    load_twiddles<G>(tws, a.tw);
    const double ky = a.kfy[iy];
    const double kz = a.kfz[zok ? zc : 0];
    const double kyz = ky * ky + kz * kz;
    for (int comp = blockIdx.z; comp < a.ncomp; comp += gridDim.z) {
        const float2* const src = a.P[comp] + tbase;
        __syncthreads();
        if (act1) {
            float2 v[G::N1];
#pragma unroll
            for (int j1 = 0; j1 < G::N1; ++j1) v[j1] = (src + (long long)(j1 * G::N2) * sxi)[voff];
            fwd_l1<G>(v, pen, tws, jj);
        }
        __syncthreads();
        float2 b[G::N2];
        if (act2) fwd_l2<G>(b, pen, jj);
        const int nops = a.Q3[comp] ? 2 : 1;  // |k| (op 1) only where its output is used
        int jl = jj;  // opaque copy: the |k| symbols are computed here, not hoisted and held
        asm volatile("" : "+v"(jl));
        if (KS && act2) {
            // F multiplies every operator: apply it to the spectrum once (one
            // more FP32 rounding than folding it into each symbol)
#pragma unroll
            for (int k2 = 0; k2 < G::N2; ++k2) {
                const double kx = a.kfx[jl + G::N1 * k2];
                const float F = kspace_factor(a, sqrtf((float)(kx * kx + kyz)));
                b[k2] = make_float2(b[k2].x * F, b[k2].y * F);
            }
        }
        // ops: 0 A1 -> Q1, 1 A3 -> Q3 (where used), 2 A2 -> Q2 (KS only)
#pragma unroll
        for (int op = 0; op < (KS ? 3 : 2); ++op) {
            if (!KS && op >= nops) break;
            if (KS && op == 1 && nops < 2) continue;
            __syncthreads();
            if (act2) {
                float2 c[G::N2];
#pragma unroll
                for (int k2 = 0; k2 < G::N2; ++k2) {
                    const double kx = a.kfx[jl + G::N1 * k2];
                    const double k2s = kx * kx + kyz;
                    // 1/sqrt(k2s): the FP32 hardware estimate (relative
                    // error below 2^-23) and one FP64 Newton step (below
                    // 2^-45), so the FP32 rounding below is that of the
                    // exact value except within 2^-45 of a rounding boundary
                    double y = __builtin_amdgcn_rsqf((float)k2s);
                    y = fma(y, fma(-0.5 * k2s * y, y, 0.5), y);
                    const double s = KS && op == 2 ? a.invN  // A2 = 1 (F already applied)
                                                   : (op == 0 ? a.invN * y : a.invN * (k2s * y));
                    float fs = (float)s;
                    // k = 0 only at element 0 (k2 = 0, jl = 0), where the
                    // estimate is inf and y NaN: both symbols are 0 there,
                    // |k|^-1 by the convention above
                    if (k2 == 0 && !(KS && op == 2) && !(k2s > 0.0)) fs = 0.f;
                    c[k2] = make_float2(b[k2].x * fs, b[k2].y * fs);
                }
                inv_la<G>(c, pen, tws + G::T2OFF, jj);
            }
            __syncthreads();
            if (act1) {
                float2 d[G::N1];
                inv_lb<G>(d, pen, jj);
                float2* const dst = (op == 0 ? a.Q1[comp] : (op == 1 ? a.Q3[comp] : a.Q2[comp])) + tbase;
#pragma unroll
                for (int j1 = 0; j1 < G::N1; ++j1) (dst + (long long)(G::N2 * j1) * sxi)[voff] = d[j1];
            }
        }
    }
}

// ------------------------------------------------------------------------
// z-direction complex transforms on rows.
//  mode 1 (inverse, stress step end): re_out + i im_out = IFFT_z(cin)   (im_out may be null)
//  mode 2 (plain, standalone transforms): cout = FFT_z^DIR(cin), in place allowed
struct ZRowsArgs {
    long long nrows;
    long long shot_stride;
    const float2* tw;
    int mode, dir, nfields;
    const float2* cin[8];
    float2* cout[8];
    float* re_out[8];
    float* im_out[8];
};

template <class G>
__global__ __launch_bounds__(kThreads) void k2_zrows(ZRowsArgs a)
{
    __shared__ float2 lds[G::P * G::PS];
    __shared__ float2 tws[G::TWN];
    constexpr int n = G::N;
    const int tid = threadIdx.x, shot = blockIdx.y;
    const int j = tid % G::N2, p1 = tid / G::N2;
    const int k1 = tid % G::N1, p2 = tid / G::N1;
    const long long r1 = (long long)blockIdx.x * G::P + p1, r2 = (long long)blockIdx.x * G::P + p2;
    const bool ok1 = p1 < G::P && r1 < a.nrows, ok2 = p2 < G::P && r2 < a.nrows;
    const long long soff = (long long)shot * a.shot_stride;
    float2* pen1 = lds + p1 * G::PS;
    float2* pen2 = lds + p2 * G::PS;
    load_twiddles<G>(tws, a.tw);
    for (int f = blockIdx.z; f < a.nfields; f += gridDim.z) {
        __syncthreads();
        if (a.mode == 2 && a.dir < 0) {
            if (ok1) {
                float2 v[G::N1];
                const float2* src = a.cin[f] + soff + r1 * n;
#pragma unroll
                for (int j1 = 0; j1 < G::N1; ++j1) v[j1] = src[j1 * G::N2 + j];
                fwd_l1<G>(v, pen1, tws, j);
            }
            __syncthreads();
            if (ok2) {
                float2 b[G::N2];
                fwd_l2<G>(b, pen2, k1);
                float2* dst = a.cout[f] + soff + r2 * n;
#pragma unroll
                for (int k2 = 0; k2 < G::N2; ++k2) dst[k1 + G::N1 * k2] = b[k2];
            }
        } else {
            if (ok2) {
                float2 b[G::N2];
                const float2* src = a.cin[f] + soff + r2 * n;
#pragma unroll
                for (int k2 = 0; k2 < G::N2; ++k2) b[k2] = src[k1 + G::N1 * k2];
                inv_la<G>(b, pen2, tws + G::T2OFF, k1);
            }
            __syncthreads();
            if (ok1) {
                float2 d[G::N1];
                inv_lb<G>(d, pen1, j);
                if (a.mode == 2) {
                    float2* dst = a.cout[f] + soff + r1 * n;
#pragma unroll
                    for (int j1 = 0; j1 < G::N1; ++j1) dst[j + G::N2 * j1] = d[j1];
                } else {
                    float* re = a.re_out[f] + soff + r1 * n;
                    float* im = a.im_out[f] ? a.im_out[f] + soff + r1 * n : nullptr;
#pragma unroll
                    for (int j1 = 0; j1 < G::N1; ++j1) {
                        re[j + G::N2 * j1] = d[j1].x;
                        if (im) im[j + G::N2 * j1] = d[j1].y;
                    }
                }
            }
        }
    }
}

// ------------------------------------------------------------------------
// 1-D shift-derivative along z of one real field (standalone entry point).
// Tile: G::P complex pencils = 2*G::P rows.
struct RowDerivArgs {
    long long nrows;
    const float2* tw;
    const float2* sym;
    const float* in;
    float* out;
};

template <class G>
__global__ __launch_bounds__(kThreads) void k2_deriv_rows(RowDerivArgs a)
{
    __shared__ float2 lds[G::P * G::PS];
    __shared__ float2 tws[G::TWN];
    constexpr int n = G::N;
    const int tid = threadIdx.x;
    const int j = tid % G::N2, p1 = tid / G::N2;
    const int k1 = tid % G::N1, p2 = tid / G::N1;
    const long long ra = (long long)blockIdx.x * 2 * G::P + 2 * p1;  // rows ra, ra+1
    const bool act1 = p1 < G::P, act2 = p2 < G::P;
    float2* pen1 = lds + p1 * G::PS;
    float2* pen2 = lds + p2 * G::PS;
    load_twiddles<G>(tws, a.tw);
    __syncthreads();
    if (act1) {
        float2 v[G::N1];
#pragma unroll
        for (int j1 = 0; j1 < G::N1; ++j1) {
            const int i = j1 * G::N2 + j;
            v[j1] = make_float2(ra < a.nrows ? a.in[ra * n + i] : 0.f, ra + 1 < a.nrows ? a.in[(ra + 1) * n + i] : 0.f);
        }
        fwd_l1<G>(v, pen1, tws, j);
    }
    __syncthreads();
    float2 b[G::N2];
    if (act2) {
        fwd_l2<G>(b, pen2, k1);
#pragma unroll
        for (int k2 = 0; k2 < G::N2; ++k2) b[k2] = c_mul(b[k2], a.sym[k1 + G::N1 * k2]);
    }
    __syncthreads();
    if (act2) inv_la<G>(b, pen2, tws + G::T2OFF, k1);
    __syncthreads();
    if (act1) {
        float2 d[G::N1];
        inv_lb<G>(d, pen1, j);
#pragma unroll
        for (int j1 = 0; j1 < G::N1; ++j1) {
            const int i = j + G::N2 * j1;
            if (ra < a.nrows) a.out[ra * n + i] = d[j1].x;
            if (ra + 1 < a.nrows) a.out[(ra + 1) * n + i] = d[j1].y;
        }
    }
}

// ------------------------------------------------------------------------
// Stress update (cuviscoelas_3D.cu:800-1015, generalised to VTI moduli) fused
// with the z-derivatives of the 15 family fields, the stress C-PML memory
// variables on the face slabs, and the source injection
// (cuviscoelas_3D.cu:771-777).  Tile: R = G::P/8 rows; 8 complex pencils per
// row carry the 15 fields (two per pencil, same z-symbol); their derivatives
// are staged in LDS, then every thread updates R*nz/256 cells.
struct StressZArgs {
    int nx, ny, nz, pml;
    int ntiles, nshots, run;  // grid: tile_shot()
    long long nrows;
    float dt;
    long long shot_stride, slabx_stride, slaby_stride, slabz_stride;
    const float2* tw;
    const float2 *symp, *symm;
    const float *pax, *pbx, *paxh, *pbxh;
    const float *pay, *pby, *payh, *pbyh;
    const float *paz, *pbz, *pazh, *pbzh;
    // [component][A1 v, A2 v, A3 v, A1 dv, A2 dv]; A2 = 1 (the fields v and
    // dv themselves) unless the k-space correction is on
    const float* fam[3][5];
    float* vel[3];  // vx vy vz: force sources are added here
    // x- and y-derivatives, blocked groups of 15 per shot:
    //   dxb: three groups of 5, dx_group floats apart: ex (dx vx-family),
    //        gx (dx vz-family), hx (dx vy-family), each [shot][N/64][5][64]
    //   dyb: ey 0-4 (dy vy-family), gy 5-9 (dy vz-family), hy 10-14 (dy vx-family)
    const float *dxb, *dyb;
    long long dx_group;
    long long blk_stride;  // per-shot field length of the blocked groups (N rounded up to 64)
    // coefficient groups G1 {P, S, S-bar} and G2 {H, D} (VTI; else null)
    const float *g1, *g2;
    float* sig[6];  // pxx pyy pzz pxy pxz pyz
    float* phx[5];
    float* phy[5];
    float* phz[4];
    const int* src_kind;
    const long long* src_ip;
    const float* rik;
    const int* it;
};

// Rows of the stress update's workgroup tile: 8 pencils (kinds) per row.  One
// pass of G::P pencils gives P/8 rows; where that leaves a third or more of
// the pencil slots unused (P = 12..15) and 2P/8 rows still fit 54 KB of LDS
// (3 workgroups per CU), the first transform level takes two passes over
// 2P/8 rows instead (n = 200 = 10 x 20: 3 rows instead of 1).
template <class G>
struct StressTile {
    static constexpr int P1 = kThreads / G::N2, P2 = kThreads / G::N1;  // pencils per pass of each level
    static constexpr int R2 = 2 * G::P / 8;
    static constexpr bool kWide = G::P / 8 == 1 && G::P >= 12 && 8L * (8 * R2 * G::PS + G::TWN) <= 54L * 1024 &&
                                  8 * R2 <= P2;
    static constexpr int R = kWide ? R2 : G::P / 8;
    static constexpr int RD1 = (8 * R + P1 - 1) / P1;  // passes of the first level
    static_assert(8 * R <= P2, "the second level takes one pass");
};

// pencil kind -> (component, family member of the real part, of the imaginary part)
// staging field index = 2*kind (+1 for the imaginary part)
//   0: A1 vx + i A1 dvx (D-)   1: A1 vy + i A1 dvy (D-)   2: A1 vz + i A1 dvz (D+)
//   3: vx + i dvx (D-)         4: vy + i dvy (D-)         5: vz + i dvz (D+)
//   6: A3 vx + i A3 vy (D-)    7: A3 vz (D+)
template <class G, bool VTI>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(QVTI_STRESSZ_WAVES))) void k2_stress_z(
    StressZArgs a)
{
    using T = StressTile<G>;
    constexpr int R = T::R, NP = 8 * R;
    constexpr int n = G::N;
    static_assert(R >= 1, "tile too small");
    static_assert(G::PS >= n + 1, "staging must fit in the pencil buffer");
    __shared__ float2 lds[NP * G::PS];
    __shared__ float2 tws[G::TWN];
    const TileShot ts = tile_shot(a.ntiles, a.nshots, a.run);
    if (ts.tile < 0) return;  // padding workgroup (whole workgroup, before any barrier)
    const int tid = threadIdx.x, shot = ts.shot;
    // level 1: N2 lanes per pencil, T::P1 pencils per pass; level 2: N1 lanes
    // per pencil, all NP pencils in one pass
    const int j = tid % G::N2, u1 = tid / G::N2;
    const int k1 = tid % G::N1, q2 = tid / G::N1;
    const bool act2 = q2 < NP;
    const long long r0 = (long long)ts.tile * R;
    const long long soff = (long long)shot * a.shot_stride;
    float2* pen2 = lds + q2 * G::PS;
    load_twiddles<G>(tws, a.tw);
    __syncthreads();
    float2 d[T::RD1][G::N1];
    {
#pragma unroll
        for (int r = 0; r < T::RD1; ++r) {
            const int q1 = u1 + r * T::P1;
            if (u1 < T::P1 && q1 < NP) {
                const long long row1 = r0 + q1 / 8;
                const bool ok1 = row1 < a.nrows;
                const float* fre;
                const float* fim;
                switch (q1 % 8) {
                case 0: fre = a.fam[0][0]; fim = a.fam[0][3]; break;
                case 1: fre = a.fam[1][0]; fim = a.fam[1][3]; break;
                case 2: fre = a.fam[2][0]; fim = a.fam[2][3]; break;
                case 3: fre = a.fam[0][1]; fim = a.fam[0][4]; break;
                case 4: fre = a.fam[1][1]; fim = a.fam[1][4]; break;
                case 5: fre = a.fam[2][1]; fim = a.fam[2][4]; break;
                case 6: fre = a.fam[0][2]; fim = a.fam[1][2]; break;
                default: fre = a.fam[2][2]; fim = nullptr; break;
                }
                float2 v[G::N1];
#pragma unroll
                for (int j1 = 0; j1 < G::N1; ++j1) {
                    const long long g = soff + row1 * n + j1 * G::N2 + j;
                    v[j1] = ok1 ? make_float2(fre[g], fim ? fim[g] : 0.f) : make_float2(0.f, 0.f);
                }
                fwd_l1<G>(v, lds + q1 * G::PS, tws, j);
            }
        }
        __syncthreads();
        float2 b[G::N2];
        if (act2) {
            fwd_l2<G>(b, pen2, k1);
            const int k2kind = q2 % 8;
            const float2* S = (k2kind == 2 || k2kind == 5 || k2kind == 7) ? a.symp : a.symm;
#pragma unroll
            for (int k2 = 0; k2 < G::N2; ++k2) b[k2] = c_mul(b[k2], S[k1 + G::N1 * k2]);
        }
        __syncthreads();
        if (act2) inv_la<G>(b, pen2, tws + G::T2OFF, k1);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < T::RD1; ++r) {
            const int q1 = u1 + r * T::P1;
            if (u1 < T::P1 && q1 < NP) inv_lb<G>(d[r], lds + q1 * G::PS, j);
        }
        __syncthreads();
        // stage: st[field][rr][iz] with rows padded to n + 1 floats (odd), so
        // the rows written by one wave fall in different banks, field = 2*kind + part
        float* st = reinterpret_cast<float*>(lds);
#pragma unroll
        for (int r = 0; r < T::RD1; ++r) {
            const int q1 = u1 + r * T::P1;
            if (u1 < T::P1 && q1 < NP) {
                const int kind = q1 % 8, rr = q1 / 8;
#pragma unroll
                for (int j1 = 0; j1 < G::N1; ++j1) {
                    const int iz = j + G::N2 * j1;
                    st[((2 * kind) * R + rr) * (n + 1) + iz] = d[r][j1].x;
                    st[((2 * kind + 1) * R + rr) * (n + 1) + iz] = d[r][j1].y;
                }
            }
        }
        __syncthreads();
    }
    const float* st = reinterpret_cast<const float*>(lds);
    const int pml = a.pml;
    const float dt = a.dt;
    const int it = *a.it;
    const int kind = a.src_kind[shot];
    const long long sip = a.src_ip[shot];
    // uniform block bases; per-cell offsets below are unsigned 32-bit
    const long long cbase = r0 * n;          // flat index of the tile's first cell
    const long long fbase = soff + cbase;    // per-shot natural fields
    const float* __restrict__ dxs = a.dxb + (long long)shot * 5 * a.blk_stride;
    const float* __restrict__ dys = a.dyb + (long long)shot * kG1 * a.blk_stride;
    constexpr bool vti = VTI;  // G2 (H, D sets) present: a compile-time choice keeps the loads branch-free
    const long long sxb = (long long)shot * a.slabx_stride, syb = (long long)shot * a.slaby_stride,
                    szb = (long long)shot * a.slabz_stride;
#pragma unroll 1
    for (int cell = tid; cell < R * n; cell += kThreads) {
        const int rr = cell / n, iz = cell - rr * n;
        const long long row = r0 + rr;
        if (row >= a.nrows) continue;
        const int ix = (int)(row / a.ny), iy = (int)(row - (long long)ix * a.ny);
        const unsigned o = (unsigned)cell;  // == rr * n + iz
        auto S = [&](int f) { return st[(f * R + rr) * (n + 1) + iz]; };
        if (ix < a.nx - 1 && iy < a.ny - 1 && iz < n - 1) {
            // staging fields of the family members j = 0..4 (A1 v, v, A3 v, A1 dv, dv)
            const int fz[5] = {4, 10, 14, 5, 11};   // vz family (D+): normal dz vz
            const int fgx[5] = {0, 6, 12, 1, 7};    // vx family (D-): dz vx for pxz
            const int fhy[5] = {2, 8, 13, 3, 9};    // vy family (D-): dz vy for pyz
            // blocked addresses of the cell: one pointer per group, the
            // fields at immediate offsets f*64 floats
            const unsigned ip = (unsigned)cbase + o;
            const unsigned q15 = (ip >> 6) * (kBlk * kG1) + (ip & 63u);
            const float* __restrict__ dxp = dxs + ((ip >> 6) * (kBlk * 5) + (ip & 63u));
            const float* __restrict__ dyp = dys + q15;
            const float* __restrict__ g1p = a.g1 + q15;
            const float* __restrict__ g2p = vti ? a.g2 + ((ip >> 6) * (kBlk * kG2) + (ip & 63u)) : nullptr;
            float A[14];
#pragma unroll
            for (int k = 0; k < 14; ++k) A[k] = 0.f;
#pragma unroll
            for (int jf = 0; jf < 5; ++jf) {
                const float ez = S(fz[jf]), gz = S(fgx[jf]), hz = S(fhy[jf]);
                const float ex = dxp[jf * kBlk], gx = dxp[a.dx_group + jf * kBlk], hx = dxp[2 * a.dx_group + jf * kBlk];
                const float ey = dyp[jf * kBlk], gy = dyp[(5 + jf) * kBlk], hy = dyp[(10 + jf) * kBlk];
                const float c33 = g1p[jf * kBlk], cs = g1p[(5 + jf) * kBlk], csb = g1p[(10 + jf) * kBlk];
                float c11 = c33, c13 = c33;
                if (vti) {
                    c11 = g2p[jf * kBlk];
                    c13 = g2p[(5 + jf) * kBlk];
                }
                const float c12 = c11 - 2 * cs, c13d = c13 - 2 * cs;
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
            // Every load of the cell is issued before its first store: the
            // stress and C-PML arrays are distinct, but the compiler cannot
            // know it, so loads placed after a store would each wait for it.
            float sg[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) sg[k] = a.sig[k][fbase + o];
            const bool xs = ix < pml || ix >= a.nx - pml;
            const bool ys = iy < pml || iy >= a.ny - pml;
            const bool zs = iz < pml || iz >= n - pml;
            unsigned ox = 0, oy = 0, oz = 0;
            float qx[5], qy[5], qz[4];
            if (xs) {
                const int si = ix < pml ? ix : ix - (a.nx - 2 * pml);
                ox = (unsigned)(((long long)si * a.ny + iy) * n + iz);
#pragma unroll
                for (int k = 0; k < 5; ++k) qx[k] = a.phx[k][sxb + ox];
            }
            if (ys) {
                const int si = iy < pml ? iy : iy - (a.ny - 2 * pml);
                oy = (unsigned)(((long long)ix * 2 * pml + si) * n + iz);
#pragma unroll
                for (int k = 0; k < 5; ++k) qy[k] = a.phy[k][syb + oy];
            }
            if (zs) {
                const int si = iz < pml ? iz : iz - (n - 2 * pml);
                oz = (unsigned)(row * 2 * pml + si);
#pragma unroll
                for (int k = 0; k < 4; ++k) qz[k] = a.phz[k][szb + oz];
            }
            float fxx_x = 0.f, fxy_x = 0.f, fxz_x = 0.f, fzx = 0.f, fyx = 0.f;
            float fyx_y = 0.f, fyy_y = 0.f, fyz_y = 0.f, fxy = 0.f, fzy = 0.f;
            float fzx_z = 0.f, fzz_z = 0.f, fxz = 0.f, fyz = 0.f;
            if (xs) {
                const float Ax = a.pax[ix], Bx = a.pbx[ix], Ah = a.paxh[ix], Bh = a.pbxh[ix];
                fxx_x = qx[0] + dt * (-Ax * qx[0] - Bx * A[0]);
                fxy_x = qx[1] + dt * (-Ax * qx[1] - Bx * A[1]);
                fxz_x = qx[2] + dt * (-Ax * qx[2] - Bx * A[2]);
                fzx = qx[3] + dt * (-Ah * qx[3] - Bh * A[9]);
                fyx = qx[4] + dt * (-Ah * qx[4] - Bh * A[11]);
                a.phx[0][sxb + ox] = fxx_x;
                a.phx[1][sxb + ox] = fxy_x;
                a.phx[2][sxb + ox] = fxz_x;
                a.phx[3][sxb + ox] = fzx;
                a.phx[4][sxb + ox] = fyx;
            }
            if (ys) {
                const float Ay = a.pay[iy], By = a.pby[iy], Ah = a.payh[iy], Bh = a.pbyh[iy];
                fyx_y = qy[0] + dt * (-Ay * qy[0] - By * A[3]);
                fyy_y = qy[1] + dt * (-Ay * qy[1] - By * A[4]);
                fyz_y = qy[2] + dt * (-Ay * qy[2] - By * A[5]);
                fxy = qy[3] + dt * (-Ah * qy[3] - Bh * A[10]);
                fzy = qy[4] + dt * (-Ah * qy[4] - Bh * A[13]);
                a.phy[0][syb + oy] = fyx_y;
                a.phy[1][syb + oy] = fyy_y;
                a.phy[2][syb + oy] = fyz_y;
                a.phy[3][syb + oy] = fxy;
                a.phy[4][syb + oy] = fzy;
            }
            if (zs) {
                const float Az = a.paz[iz], Bz = a.pbz[iz], Ah = a.pazh[iz], Bh = a.pbzh[iz];
                fzx_z = qz[0] + dt * (-Az * qz[0] - Bz * A[6]);
                fzz_z = qz[1] + dt * (-Az * qz[1] - Bz * A[7]);
                fxz = qz[2] + dt * (-Ah * qz[2] - Bh * A[8]);
                fyz = qz[3] + dt * (-Ah * qz[3] - Bh * A[12]);
                a.phz[0][szb + oz] = fzx_z;
                a.phz[1][szb + oz] = fzz_z;
                a.phz[2][szb + oz] = fxz;
                a.phz[3][szb + oz] = fyz;
            }
            a.sig[0][fbase + o] = sg[0] + dt * (A[0] + A[3] + A[6] + (fxx_x + fyx_y + fzx_z));
            a.sig[1][fbase + o] = sg[1] + dt * (A[1] + A[4] + A[6] + (fxy_x + fyy_y + fzx_z));
            a.sig[2][fbase + o] = sg[2] + dt * (A[2] + A[5] + A[7] + (fxz_x + fyz_y + fzz_z));
            a.sig[3][fbase + o] = sg[3] + dt * (A[10] + fxy + A[11] + fyx);
            a.sig[4][fbase + o] = sg[4] + dt * (A[8] + fxz + A[9] + fzx);
            a.sig[5][fbase + o] = sg[5] + dt * (A[12] + fyz + A[13] + fzy);
        }
        if (row * n + iz == sip) {
            // source after the stress update; no other workgroup of this
            // launch touches this row of v (at most this one read it, for its
            // z-derivative, staged above), so the write cannot race.
            const float s = a.rik[it];
            const long long g = fbase + o;
            if (kind == 3) {
                a.sig[0][g] += s;
                a.sig[1][g] += s;
                a.sig[2][g] += s;
            } else if (kind >= 0 && kind < 3) {
                a.vel[kind][g] += s;
            }
        }
    }
}

}  // namespace qvti
