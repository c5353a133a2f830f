// qvti_fft2l.h -- two-level register FFT for gfx950 pencils.
//
// A length n = N1*N2 transform is split Cooley-Tukey style into N2-point and
// N1-point DFTs that each thread computes entirely in registers:
//
//   forward level 1 (thread = pencil p, column j2):  a[j1] = x[j1*N2 + j2],
//       a = DFT_N1(a), a[k1] *= W_n^(j2*k1), stored to LDS as Y[p][k1][j2]
//   forward level 2 (thread = pencil p, row k1):     b[j2] = Y[p][k1][j2],
//       b = DFT_N2(b)  ->  b[k2] = X[k1 + N1*k2]
//   inverse level A (same thread as level 2):        b = IDFT_N2(b),
//       b[ja] *= W_n^(-ja*k1), stored as Z[p][ja][k1]
//   inverse level B (same thread as level 1):        d[k1] = Z[p][ja][k1],
//       d = IDFT_N1(d)  ->  d[jb] = x[ja + N2*jb]   (unnormalised)
//
// so a forward transform, a pointwise spectral operator and the inverse
// transform cost two LDS exchanges, and each thread owns the same N1 cells
// x[j2 + N2*j1] for the load and for the store.  The in-register DFTs are
// compile-time mixed-radix (4, 2, 3, 5) Cooley-Tukey with constant twiddles;
// the inter-level twiddles W_n^e come from an LDS table built on the host in
// FP64 and rounded to FP32.  This replaces the reference's cuFFT C2C plans
// (cuviscoelas_3D.cu:1703-1704).
#pragma once
#include <hip/hip_runtime.h>

namespace qvti {

// ---- compile-time trigonometry (Taylor series in FP64 after reduction to
// [-pi/4, pi/4]; exact to ~1e-16, then rounded to FP32)
constexpr double kCxPi = 3.14159265358979323846;

constexpr double cx_taylor_sin(double x)
{
#pragma clang fp contract(off)
    double term = x, sum = x;
    for (int k = 1; k < 14; ++k) {
        term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double cx_taylor_cos(double x)
{
#pragma clang fp contract(off)
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 14; ++k) {
        term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
        sum += term;
    }
    return sum;
}
// cos and sin of 2*pi*e/n for integers e, n (exact octant reduction)
constexpr double cx_cos2pi(long e, long n)
{
#pragma clang fp contract(off)
    e %= n;
    if (e < 0) e += n;
    // angle = 2 pi e / n; use 8*e/n octant
    const long o8 = (8 * e) / n;           // octant 0..7
    const long rem8 = 8 * e - o8 * n;      // 0 <= rem8 < n, angle within octant = (pi/4) * rem8 / n
    const double t = (kCxPi / 4.0) * (double)rem8 / (double)n;
    const double c = cx_taylor_cos(t), s = cx_taylor_sin(t);
    const double r2 = 0.70710678118654752440;
    switch (o8) {  // angle = o8*pi/4 + t
    case 0: return c;
    case 1: return r2 * (c - s);
    case 2: return -s;
    case 3: return -r2 * (c + s);
    case 4: return -c;
    case 5: return -r2 * (c - s);
    case 6: return s;
    default: return r2 * (c + s);
    }
}
constexpr double cx_sin2pi(long e, long n)
{
    // sin(a) = cos(a - pi/2) = cos(2 pi (e - n/4) / n); use 4e - n over 4n
    return cx_cos2pi(4 * e - n, 4 * n);
}

// W_n^e = exp(-2 pi i e / n) rounded to FP32 from the FP64 series above.
// With contraction off the host and the device evaluate the same IEEE
// operations, so a table built on either side has the same bits.
__host__ __device__ inline float2 twiddle_w(long e, long n)
{
    return make_float2((float)cx_cos2pi(e, n), (float)(-cx_sin2pi(e, n)));
}

// exp(-2 pi i e / N) as two FP32 constants
template <int N>
struct CTw {
    float re[N], im[N];
    constexpr CTw() : re(), im()
    {
        for (int e = 0; e < N; ++e) {
            re[e] = (float)cx_cos2pi(e, N);
            im[e] = (float)(-cx_sin2pi(e, N));
        }
    }
};
template <int N>
struct TwTab {
    static constexpr CTw<N> t{};
};

__host__ __device__ __forceinline__ float2 c_mul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ __forceinline__ float2 c_add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ float2 c_sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by S*i
template <int S>
__host__ __device__ __forceinline__ float2 c_muli(float2 a)
{
    return S < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

// multiply by the constant exp(S 2 pi i e / N), e compile-time after unrolling
template <int N, int S>
__host__ __device__ __forceinline__ float2 c_twc(float2 a, int e)
{
    const int r = e % N;
    if (r == 0) return a;
    // quarter turns are exact swaps and negations
    if (4 * r == N) return c_muli<S>(a);
    if (2 * r == N) return make_float2(-a.x, -a.y);
    if (4 * r == 3 * N) return c_muli<-S>(a);
    const float wr = TwTab<N>::t.re[r];
    const float wi = S < 0 ? TwTab<N>::t.im[r] : -TwTab<N>::t.im[r];
    return make_float2(a.x * wr - a.y * wi, a.x * wi + a.y * wr);
}

constexpr int first_radix(int n)
{
    return n % 4 == 0 && n != 8 ? 4 : (n % 2 == 0 ? 2 : (n % 3 == 0 ? 3 : (n % 5 == 0 ? 5 : n)));
}

// In-register DFT of N points stored at v[0], v[ST], ..., v[(N-1)*ST]:
// X[k] = sum_j x[j] exp(S 2 pi i j k / N).
template <int N, int S, int ST = 1>
__host__ __device__ __forceinline__ void rdft(float2* v)
{
    if constexpr (N == 1) {
        return;
    } else if constexpr (N == 2) {
        const float2 a = v[0];
        v[0] = c_add(a, v[ST]);
        v[ST] = c_sub(a, v[ST]);
    } else if constexpr (N == 3) {
        const float h = 0.86602540378443864676f;
        const float2 t1 = c_add(v[ST], v[2 * ST]);
        const float2 t2 = make_float2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
        const float2 d = c_sub(v[ST], v[2 * ST]);
        const float2 t3 = c_muli<S>(make_float2(h * d.x, h * d.y));
        v[0] = c_add(v[0], t1);
        v[ST] = c_add(t2, t3);
        v[2 * ST] = c_sub(t2, t3);
    } else if constexpr (N == 4) {
        const float2 t0 = c_add(v[0], v[2 * ST]), t1 = c_sub(v[0], v[2 * ST]);
        const float2 t2 = c_add(v[ST], v[3 * ST]), t3 = c_muli<S>(c_sub(v[ST], v[3 * ST]));
        v[0] = c_add(t0, t2);
        v[2 * ST] = c_sub(t0, t2);
        v[ST] = c_add(t1, t3);
        v[3 * ST] = c_sub(t1, t3);
    } else if constexpr (N == 5) {
        const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
        const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
        const float2 t1 = c_add(v[ST], v[4 * ST]), t2 = c_add(v[2 * ST], v[3 * ST]);
        const float2 t3 = c_sub(v[ST], v[4 * ST]), t4 = c_sub(v[2 * ST], v[3 * ST]);
        const float2 a1 = make_float2(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
        const float2 a2 = make_float2(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
        const float2 b1 = c_muli<S>(make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
        const float2 b2 = c_muli<S>(make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
        v[0] = make_float2(v[0].x + t1.x + t2.x, v[0].y + t1.y + t2.y);
        v[ST] = c_add(a1, b1);
        v[4 * ST] = c_sub(a1, b1);
        v[2 * ST] = c_add(a2, b2);
        v[3 * ST] = c_sub(a2, b2);
    } else if constexpr (N == 8) {
        // radix 2 x 4 with the constant eighth-root twiddles
        float2 e[4], o[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) { e[i] = v[2 * i * ST]; o[i] = v[(2 * i + 1) * ST]; }
        rdft<4, S, 1>(e);
        rdft<4, S, 1>(o);
        const float r = 0.70710678118654752440f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            float2 t = o[k];
            if (k == 1) t = make_float2(r * (t.x - (float)S * t.y), r * (t.y + (float)S * t.x));
            else if (k == 2) t = c_muli<S>(t);
            else if (k == 3) t = make_float2(r * (-t.x - (float)S * t.y), r * (-t.y + (float)S * t.x));
            v[k * ST] = c_add(e[k], t);
            v[(k + 4) * ST] = c_sub(e[k], t);
        }
    } else {
        // N = R*M: gather r-th decimated subsequences, M-point DFTs, twiddle,
        // R-point DFTs across them.  Output X[k1 + M*k2].
        constexpr int R = first_radix(N);
        constexpr int M = N / R;
        static_assert(R * M == N && R > 1 && R <= 5, "unsupported DFT length");
        float2 t[R][M];
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int m = 0; m < M; ++m) t[r][m] = v[(R * m + r) * ST];
#pragma unroll
        for (int r = 0; r < R; ++r) rdft<M, S, 1>(t[r]);
#pragma unroll
        for (int k1 = 0; k1 < M; ++k1) {
            float2 u[R];
#pragma unroll
            for (int r = 0; r < R; ++r) u[r] = c_twc<N, S>(t[r][k1], r * k1);
            rdft<R, S, 1>(u);
#pragma unroll
            for (int k2 = 0; k2 < R; ++k2) v[(k1 + M * k2) * ST] = u[k2];
        }
    }
}

// Per-length constants of the two-level scheme.  PW widens the tile of the
// strided-axis kernels: a workgroup of T = 256*PW threads holds P = T/NMAX
// pencils, so each row of the tile is P (complex) or 2P (real) contiguous
// z-columns wide.
template <int N1_, int N2_, int PW_ = 1>
struct L2 {
    static constexpr int N1 = N1_, N2 = N2_, N = N1_ * N2_;
    static constexpr int NMAX = N1 > N2 ? N1 : N2;
    static constexpr int PW = PW_;
    static constexpr int T = 256 * PW_;           // threads per workgroup
    static constexpr int P = T / NMAX;            // pencils per tile
    // LDS layout of one pencil: Y rows k1 of length N2 padded by one, Z rows ja
    // of length N1 padded by one; the pencil stride is odd (bank spread).
    static constexpr int YS = N2 + 1, ZS = N1 + 1;
    static constexpr int PS0 = (N1 * YS > N2 * ZS ? N1 * YS : N2 * ZS);
    static constexpr int PS = PS0 | 1;
    // Twiddle tables in LDS: T1[k1*N2 + j2] = W_N^(j2*k1) (forward level 1)
    // and T2[ja*N1 + k1] = W_N^(ja*k1) (inverse level A, conjugated on use).
    // Each thread reads its table with one base register and immediate
    // offsets.  For N1 == N2 the two tables coincide.
    static constexpr int T2OFF = N1 == N2 ? 0 : N;
    static constexpr int TWN = N + T2OFF;
    // LDS of a strided tile (pencils + twiddles) within the 160 KiB of a CU
    static constexpr bool kFits = 8L * (P * PS + TWN) <= 160L * 1024;
};

// Fill the twiddle tables (G::TWN entries) from tw[m] = W_N^m.
template <class G>
__host__ __device__ __forceinline__ void twiddle_entry(float2* t, const float2* tw, int idx)
{
    if (idx < G::N) {
        const int k1 = idx / G::N2, j2 = idx - k1 * G::N2;
        t[idx] = tw[j2 * k1];
    } else {
        const int r = idx - G::N, ja = r / G::N1, k1 = r - ja * G::N1;
        t[idx] = tw[ja * k1];
    }
}

// Forward level 1: a[j1] = x[j1*N2 + j2] -> DFT_N1 -> twiddle -> Y[p][k1][j2]
// (tw: table T1)
template <class G>
__host__ __device__ __forceinline__ void fwd_l1(float2 (&a)[G::N1], float2* pen, const float2* tw, int j2)
{
    rdft<G::N1, -1, 1>(a);
#pragma unroll
    for (int k1 = 0; k1 < G::N1; ++k1) {
        const float2 w = k1 ? tw[k1 * G::N2 + j2] : make_float2(1.f, 0.f);
        pen[k1 * G::YS + j2] = k1 ? c_mul(a[k1], w) : a[k1];
    }
}

// Forward level 2: b[j2] = Y[p][k1][j2] -> DFT_N2 -> b[k2] = X[k1 + N1*k2]
template <class G>
__host__ __device__ __forceinline__ void fwd_l2(float2 (&b)[G::N2], const float2* pen, int k1)
{
#pragma unroll
    for (int j2 = 0; j2 < G::N2; ++j2) b[j2] = pen[k1 * G::YS + j2];
    rdft<G::N2, -1, 1>(b);
}

// Inverse level A: IDFT_N2 over k2, twiddle conj(W_n^(ja*k1)) -> Z[p][ja][k1]
// (tw: table T2)
template <class G>
__host__ __device__ __forceinline__ void inv_la(float2 (&b)[G::N2], float2* pen, const float2* tw, int k1)
{
    rdft<G::N2, 1, 1>(b);
#pragma unroll
    for (int ja = 0; ja < G::N2; ++ja) {
        float2 w = tw[ja * G::N1 + k1];
        w.y = -w.y;
        pen[ja * G::ZS + k1] = (ja && k1) ? c_mul(b[ja], w) : b[ja];
    }
}

// Inverse level B: d[k1] = Z[p][ja][k1] -> IDFT_N1 -> d[jb] = x[ja + N2*jb]
template <class G>
__host__ __device__ __forceinline__ void inv_lb(float2 (&d)[G::N1], const float2* pen, int ja)
{
#pragma unroll
    for (int k1 = 0; k1 < G::N1; ++k1) d[k1] = pen[ja * G::ZS + k1];
    rdft<G::N1, 1, 1>(d);
}

}  // namespace qvti
