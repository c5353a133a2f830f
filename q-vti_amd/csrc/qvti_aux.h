// qvti_aux.h -- non-transform kernels of the engine: receivers, step
// counter, snapshots, blow-up probe and the once-per-model setup kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "qvti_layout.h"

namespace qvti {

// Receivers: rec[it][r] = v[cell_r] for the three components (after the
// source injection), cuviscoelas_3D.cu:782-792.  The step counter `it`
// advances in the first kernel of the next step (k2_deriv_strided's tick).
struct RecArgs {
    const float *vx, *vy, *vz;
    long long shot_stride;
    const long long* rec_ip;  // all shots' receivers back to back
    const int* rec_off;       // per shot: first receiver, [nshots+1]
    float *rx, *ry, *rz;      // records, per shot offset rec_off[s]*nt
    int nt;
    const int* it;
};

static __global__ __launch_bounds__(256) void k_record(RecArgs a)
{
    const int shot = blockIdx.y;
    const int r0 = a.rec_off[shot], nr = a.rec_off[shot + 1] - r0;
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= nr) return;
    const int it = *a.it;
    const long long ip = a.rec_ip[r0 + r] + (long long)shot * a.shot_stride;
    const long long o = (long long)r0 * a.nt + (long long)it * nr + r;
    a.rx[o] = a.vx[ip];
    a.ry[o] = a.vy[ip];
    a.rz[o] = a.vz[ip];
}


// Snapshot of one field over the box [lo, hi) (ix slowest) when
// it % every == 0; snapshot k lands at out[k * W], W the box's cell count.
struct SnapArgs {
    const float* f;
    int ny, nz;
    int lo[3], hi[3];
    int every;
    const int* it;
    float* out;
    long long cap;
};

static __global__ __launch_bounds__(256) void k_snapshot(SnapArgs a)
{
    const int t = *a.it;
    if (a.every <= 0 || t % a.every != 0) return;
    const long long k = t / a.every;
    const int wx = a.hi[0] - a.lo[0], wy = a.hi[1] - a.lo[1], wz = a.hi[2] - a.lo[2];
    const long long W = (long long)wx * wy * wz;
    if ((k + 1) * W > a.cap) return;
    for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < W; q += (long long)gridDim.x * 256) {
        const int ix = (int)(q / ((long long)wy * wz));
        const long long rem = q - (long long)ix * wy * wz;
        const int iy = (int)(rem / wz), iz = (int)(rem - (long long)iy * wz);
        a.out[k * W + q] = a.f[((long long)(ix + a.lo[0]) * a.ny + iy + a.lo[1]) * a.nz + iz + a.lo[2]];
    }
}

// Blow-up probe: max |v| over a field (bit pattern of a non-negative float
// orders like an int; NaN sorts above Inf).
static __global__ __launch_bounds__(256) void k_probe(const float* f, long long N, int* out)
{
    int m = 0;
    for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < N; q += (long long)gridDim.x * 256) {
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
// media (qvti_set_media, the reference's media_type, viscoelas_3D.cpp:57-60):
// 0 viscoelastic (as above); 1 dispersion only (tau terms dropped); 2
// lossless (gamma = 0); 3 Q-compensating (tau terms negated: the amplitude
// loss becomes a gain of the same size, the dispersion is kept).
__device__ __forceinline__ void q_coeffs(double rho, double v, double g, double w0, double dt, int media, double* o)
{
    const double pi = 3.1415926;
    if (media == 2) g = 0.0;
    const double c = v * cos(g * pi / 2.0);
    o[0] = -rho * c * g * w0;
    o[1] = rho * (c * c);
    o[2] = rho * pow(c, 3.0) * g / w0;
    o[3] = rho * c * g * pi / dt;
    o[4] = rho * (c * c) * (g * g) * pi / w0 / dt;
    if (media == 1) o[3] = o[4] = 0.0;
    if (media == 3) {
        o[3] = -o[3];
        o[4] = -o[4];
    }
}

// (blocked field groups and coefficient group sizes: qvti_layout.h)

struct CoefArgs {
    long long N;
    int vti, media;
    double w0, dt;
    const float *vp, *gp, *vs, *gs, *rho, *eps, *del;
    float* g1;  // blocked [N/64][15][64]: P, S (S-bar filled by k_averages)
    float* g2;  // blocked [N/64][10][64]: H, D (VTI only)
};

static __global__ __launch_bounds__(256) void k_coefficients(CoefArgs a)
{
    for (long long ip = (long long)blockIdx.x * 256 + threadIdx.x; ip < a.N; ip += (long long)gridDim.x * 256) {
        double o[5];
        const double r = a.rho[ip];
        q_coeffs(r, a.vp[ip], a.gp[ip], a.w0, a.dt, a.media, o);
        for (int j = 0; j < 5; ++j) a.g1[blk_index(ip, kG1, j)] = (float)o[j];
        q_coeffs(r, a.vs[ip], a.gs[ip], a.w0, a.dt, a.media, o);
        for (int j = 0; j < 5; ++j) a.g1[blk_index(ip, kG1, 5 + j)] = (float)o[j];
        if (a.vti) {
            const double e = a.eps[ip], dl = a.del[ip], vp = a.vp[ip], vs = a.vs[ip];
            const double vh = e == 0.0 ? vp : vp * sqrt(1.0 + 2.0 * e);
            q_coeffs(r, vh, a.gp[ip], a.w0, a.dt, a.media, o);
            for (int j = 0; j < 5; ++j) a.g2[blk_index(ip, kG2, j)] = (float)o[j];
            double vd = vp;
            if (dl != 0.0) {
                const double c33 = r * vp * vp, c44 = r * vs * vs;
                const double prod = (c33 - c44) * (c33 * (1.0 + 2.0 * dl) - c44);
                vd = sqrt((sqrt(prod > 0.0 ? prod : 0.0) + c44) / r);
            }
            q_coeffs(r, vd, a.gp[ip], a.w0, a.dt, a.media, o);
            for (int j = 0; j < 5; ++j) a.g2[blk_index(ip, kG2, 5 + j)] = (float)o[j];
        }
    }
}

// 8-corner cube average of the S set (cuviscoelas_3D.cu:903-907) and the
// buoyancies dt / (0.5 (rho + rho_next)) (cuviscoelas_3D.cu:1077-1092).
struct AvgArgs {
    int nx, ny, nz;
    long long sx, sy, N;
    float dt;
    float* g1;       // coefficient group 1: reads S (fields 5-9), writes S-bar (10-14)
    const float* rho;
    float* buoy;     // [3][N]
};

static __global__ __launch_bounds__(256) void k_averages(AvgArgs a)
{
    for (long long ip = (long long)blockIdx.x * 256 + threadIdx.x; ip < a.N; ip += (long long)gridDim.x * 256) {
        const int ix = (int)(ip / a.sx);
        const int iy = (int)((ip - (long long)ix * a.sx) / a.sy);
        const int iz = (int)(ip - (long long)ix * a.sx - (long long)iy * a.sy);
        const bool in = ix < a.nx - 1 && iy < a.ny - 1 && iz < a.nz - 1;
        for (int j = 0; j < 5; ++j) {
            auto c = [&](long long q) { return a.g1[blk_index(q, kG1, 5 + j)]; };
            float v = 0.f;
            if (in) {
                const float s8 = c(ip) + c(ip + 1) + c(ip + a.sy) + c(ip + 1 + a.sy) + c(ip + a.sx) + c(ip + a.sx + 1)
                               + c(ip + a.sx + a.sy) + c(ip + a.sx + 1 + a.sy);
                v = 0.125f * s8;
            }
            a.g1[blk_index(ip, kG1, 10 + j)] = v;
        }
        // rho + rho' is an FP32 sum in the reference, the rest FP64
        const float r = a.rho[ip];
        a.buoy[ip] = ix < a.nx - 1 ? (float)((double)a.dt / (0.5 * (double)(r + a.rho[ip + a.sx]))) : 0.f;
        a.buoy[a.N + ip] = iy < a.ny - 1 ? (float)((double)a.dt / (0.5 * (double)(r + a.rho[ip + a.sy]))) : 0.f;
        a.buoy[2 * a.N + ip] = iz < a.nz - 1 ? (float)((double)a.dt / (0.5 * (double)(r + a.rho[ip + 1]))) : 0.f;
    }
}

}  // namespace qvti
