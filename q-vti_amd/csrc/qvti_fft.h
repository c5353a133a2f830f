// qvti_fft.h -- LDS-staged mixed-radix (2/3/4/5) Stockham FFT for gfx950.
//
// A workgroup holds a tile of P complex pencils of length n in LDS (pencil
// stride n+1 float2 so that the column-wise tile loads of strided axes hit
// distinct banks).  Each radix pass reads every butterfly's inputs into
// registers, barriers, writes the outputs to their autosort positions and
// barriers again, so one LDS buffer suffices (in-place Stockham).  Twiddles
// exp(-2 pi i t / n) come from an LDS copy of a table built on the host in
// FP64 and rounded to FP32.
//
// The transform replaces the reference's cuFFT C2C plans
// (cuviscoelas_3D.cu:1703-1704); it is unnormalised like cuFFT.
#pragma once
#include <hip/hip_runtime.h>

namespace qvti {

constexpr int kThreads = 256;       // threads per workgroup of every pencil kernel
constexpr int kTileMax = 4096;      // complex values per LDS tile (32 KiB)
constexpr int kMaxPasses = 16;

struct FftPlan {
    int n;
    int npass;
    int radix[kMaxPasses];
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by s*i (s = +1 or -1)
__device__ __forceinline__ float2 cmuli(float2 a, float s) { return make_float2(-s * a.y, s * a.x); }

// In-register DFT of size R, X[q] = sum_r v[r] exp(s 2 pi i q r / R), s = -1 forward.
template <int R>
__device__ __forceinline__ void butterfly(float2 (&v)[R], float s);

template <>
__device__ __forceinline__ void butterfly<2>(float2 (&v)[2], float)
{
    float2 a = v[0];
    v[0] = cadd(a, v[1]);
    v[1] = csub(a, v[1]);
}

template <>
__device__ __forceinline__ void butterfly<3>(float2 (&v)[3], float s)
{
    const float h = 0.86602540378443864676f;  // sin(2 pi / 3)
    float2 t1 = cadd(v[1], v[2]);
    float2 t2 = make_float2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
    float2 t3 = cmuli(csub(v[1], v[2]), s * h);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, t3);
    v[2] = csub(t2, t3);
}

template <>
__device__ __forceinline__ void butterfly<4>(float2 (&v)[4], float s)
{
    float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    float2 t2 = cadd(v[1], v[3]), t3 = cmuli(csub(v[1], v[3]), s);
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
}

template <>
__device__ __forceinline__ void butterfly<5>(float2 (&v)[5], float s)
{
    const float c1 = 0.30901699437494742410f;   // cos(2 pi / 5)
    const float c2 = -0.80901699437494742410f;  // cos(4 pi / 5)
    const float s1 = 0.95105651629515357212f;   // sin(2 pi / 5)
    const float s2 = 0.58778525229247312917f;   // sin(4 pi / 5)
    float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    float2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    float2 a1 = make_float2(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
    float2 a2 = make_float2(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
    float2 b1 = cmuli(make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y), s);
    float2 b2 = cmuli(make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y), s);
    v[0] = make_float2(v[0].x + t1.x + t2.x, v[0].y + t1.y + t2.y);
    v[1] = cadd(a1, b1);
    v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2);
    v[3] = csub(a2, b2);
}

// One Stockham pass of radix R over P pencils (pencil stride ps) in LDS.
// s = -1 forward, +1 inverse.  Butterfly b of the tile maps to pencil b/m and
// index j = b%m within it; inputs j + r*m, outputs (j/Ns)*Ns*R + j%Ns + r*Ns.
template <int R>
__device__ __forceinline__ void stockham_pass(float2* buf, const float2* tw, int n, int P, int ps, int Ns, float s)
{
    constexpr int MAXB = (kTileMax + R * kThreads - 1) / (R * kThreads);
    const int m = n / R;
    const int ts = n / (Ns * R);
    const int nb = P * m;
    float2 v[MAXB][R];
    int dst[MAXB];
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int b = threadIdx.x + q * kThreads;
        dst[q] = -1;
        if (b < nb) {
            const int pen = b / m;
            const int j = b - pen * m;
            const int k = j % Ns;
            const float2* src = buf + pen * ps;
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = src[j + r * m];
#pragma unroll
            for (int r = 1; r < R; ++r) {
                float2 w = tw[k * r * ts];
                w.y *= -s;  // table holds exp(-i..): forward (s=-1) keeps it, inverse conjugates
                v[q][r] = cmul(v[q][r], w);
            }
            butterfly<R>(v[q], s);
            dst[q] = pen * ps + (j - k) * R + k;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        if (dst[q] >= 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) buf[dst[q] + r * Ns] = v[q][r];
        }
    }
    __syncthreads();
}

// Full transform of the tile.  Caller guarantees P*n <= kTileMax, a barrier
// before the call (tile and twiddles written) and gets one after.
__device__ __forceinline__ void fft_tile(float2* buf, const float2* tw, const FftPlan& pl, int P, int ps, float s)
{
    int Ns = 1;
    for (int p = 0; p < pl.npass; ++p) {
        const int R = pl.radix[p];
        if (R == 4) stockham_pass<4>(buf, tw, pl.n, P, ps, Ns, s);
        else if (R == 2) stockham_pass<2>(buf, tw, pl.n, P, ps, Ns, s);
        else if (R == 3) stockham_pass<3>(buf, tw, pl.n, P, ps, Ns, s);
        else stockham_pass<5>(buf, tw, pl.n, P, ps, Ns, s);
        Ns *= R;
    }
}

}  // namespace qvti
