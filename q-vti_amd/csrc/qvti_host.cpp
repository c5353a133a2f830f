// qvti_host.cpp -- host-side helpers of the engine's C ABI: the reference
// host's wavelet, gamma and receiver geometry (viscoelas_3D.cpp) and seeded
// synthetic models for the benchmark configurations.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <thread>
#include <vector>

#include "qvti.h"
#include "qvti_internal.h"

namespace {

const double kRefPi = 3.1415926;  // viscoelas_3D.cpp:14

// splitmix64 seeding a xoshiro256** stream; u = (x >> 11) * 2^-53.
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed)
    {
        for (int i = 0; i < 4; ++i) {
            seed += 0x9E3779B97F4A7C15ULL;
            uint64_t z = seed;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
            s[i] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal()
    {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
    }
};

template <class F>
void parallel_for(long long n, F f)
{
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([=] {
            for (long long i = t; i < n; i += nt) f(i);
        });
    for (auto& x : th) x.join();
}

// Separable Gaussian filter (reflecting edges), radius 3 sigma, in place.
void gauss3d(std::vector<double>& a, int nx, int ny, int nz, double sigma)
{
    const int R = std::max(1, (int)ceil(3.0 * sigma));
    std::vector<double> w(2 * R + 1);
    double sw = 0;
    for (int k = -R; k <= R; ++k) sw += w[k + R] = exp(-0.5 * k * k / (sigma * sigma));
    for (auto& x : w) x /= sw;
    auto refl = [](int i, int n) {
        while (i < 0 || i >= n) i = i < 0 ? -i - 1 : 2 * n - i - 1;
        return i;
    };
    const int dims[3] = {nx, ny, nz};
    const long long str[3] = {(long long)ny * nz, nz, 1};
    std::vector<double> out(a.size());
    for (int ax = 0; ax < 3; ++ax) {
        const int n = dims[ax];
        const long long npen = (long long)nx * ny * nz / n;
        parallel_for(npen, [&, ax, n](long long p) {
            long long base;
            if (ax == 2) base = p * nz;
            else if (ax == 1) base = (p / nz) * (long long)ny * nz + (p % nz);
            else base = p;
            for (int i = 0; i < n; ++i) {
                double acc = 0;
                for (int k = -R; k <= R; ++k) acc += w[k + R] * a[base + refl(i + k, n) * str[ax]];
                out[base + i * str[ax]] = acc;
            }
        });
        a.swap(out);
    }
}

}  // namespace

extern "C" int qvti_ricker(int nt, float dt, float f0, float* rik)
{
    if (nt < 0 || !rik || !(f0 > 0)) return qvti_fail(QVTI_E_ARG, "bad Ricker arguments");
    // viscoelas_3D.cpp:498-502: tmp is an FP32 variable; exp(-tmp) takes the
    // float overload of the C++ math header.
    for (int it = 0; it < nt; ++it) {
        const float tmp = (float)pow(kRefPi * f0 * (it * dt - 1 / f0), 2.0);
        rik[it] = (float)((1.0 - 2.0 * tmp) * (double)expf(-tmp));
    }
    return QVTI_OK;
}

extern "C" int qvti_gamma_from_q(const float* q, float* gamma, long long n)
{
    if (!q || !gamma || n < 0) return qvti_fail(QVTI_E_ARG, "bad arguments");
    for (long long i = 0; i < n; ++i) gamma[i] = (float)(atan(1.0 / q[i]) / kRefPi);  // viscoelas_3D.cpp:287-288
    return QVTI_OK;
}

extern "C" int qvti_reference_receivers(int pml, int rnx_max, int rny_max, int dr, int* r_ix, int* r_iy)
{
    if (!r_ix || !r_iy || rnx_max < 0 || rny_max < 1) return qvti_fail(QVTI_E_ARG, "bad arguments");
    const int R = rnx_max * rny_max;
    for (int ip = 0; ip < R; ++ip) {  // viscoelas_3D.cpp:181-185
        r_ix[ip] = pml + ip / rny_max * dr;
        r_iy[ip] = pml + ip % rny_max * dr;
    }
    return QVTI_OK;
}

extern "C" int qvti_make_model(int kind, uint64_t seed, int nx, int ny, int nz, int nlay, const double* p,
                               float* vp, float* vs, float* rho, float* qp, float* qs, float* eps, float* delta)
{
    if (!p || !vp || !vs || !rho || !qp || !qs || nx < 1 || ny < 1 || nz < 1)
        return qvti_fail(QVTI_E_ARG, "bad model arguments");
    const long long N = (long long)nx * ny * nz;
    Rng rng(seed);
    auto gardner = [](double v) { return 310.0 * pow(v, 0.25); };
    if (kind == 0) {
        for (long long i = 0; i < N; ++i) {
            vp[i] = (float)p[0];
            vs[i] = (float)p[1];
            rho[i] = (float)p[2];
            qp[i] = (float)p[3];
            qs[i] = (float)p[4];
            if (eps) eps[i] = 0.f;
            if (delta) delta[i] = 0.f;
        }
        return QVTI_OK;
    }
    if (kind == 1) {
        if (nlay < 1) return qvti_fail(QVTI_E_ARG, "nlay must be >= 1");
        std::vector<double> lv(nlay), le(nlay), ld(nlay), lq(nlay);
        for (int l = 0; l < nlay; ++l) {
            lv[l] = p[0] + (p[1] - p[0]) * rng.uniform();
            le[l] = p[2] * rng.uniform();
            ld[l] = le[l] * rng.uniform();
            lq[l] = p[3] + (p[4] - p[3]) * rng.uniform();
        }
        for (long long i = 0; i < N; ++i) {
            const int iz = (int)(i % nz);
            const int l = std::min(nlay - 1, (int)((long long)iz * nlay / nz));
            vp[i] = (float)lv[l];
            vs[i] = (float)(lv[l] / 1.73);
            rho[i] = (float)gardner(lv[l]);
            qp[i] = (float)lq[l];
            qs[i] = (float)(lq[l] / 2.0);
            if (eps) eps[i] = (float)le[l];
            if (delta) delta[i] = (float)ld[l];
        }
        return QVTI_OK;
    }
    if (kind == 2) {
        std::vector<double> a(N);
        for (long long i = 0; i < N; ++i) a[i] = rng.normal();
        gauss3d(a, nx, ny, nz, p[5] > 0 ? p[5] : 15.0);
        double lo = a[0], hi = a[0];
        for (double x : a) { lo = std::min(lo, x); hi = std::max(hi, x); }
        const double span = hi > lo ? hi - lo : 1.0;
        const double emin = p[6];
        for (long long i = 0; i < N; ++i) {
            const double t = (a[i] - lo) / span;  // 0 .. 1
            const double v = p[0] + (p[1] - p[0]) * t;
            vp[i] = (float)v;
            vs[i] = (float)(v / 1.73);
            rho[i] = (float)gardner(v);
            const double q = p[3] + (p[4] - p[3]) * t;
            qp[i] = (float)q;
            qs[i] = (float)(q / 2.0);
            const double e = emin + (p[2] - emin) * (1.0 - t);
            if (eps) eps[i] = (float)e;
            if (delta) delta[i] = (float)(e / 2.0);
        }
        return QVTI_OK;
    }
    return qvti_fail(QVTI_E_ARG, "unknown model kind %d", kind);
}

// ---- data formats on either side of the path --------------------------------

// Edge replication of an interior model into the total grid (C-PML frame):
// viscoelas_3D.cpp:576-655 copies the first/last interior plane outward along
// z, then y, then x, which equals reading the interior at the clamped index.
extern "C" int qvti_pad_model(const float* interior, int ntx, int nty, int ntz, int pml, float* total)
{
    if (!interior || !total) return qvti_fail(QVTI_E_ARG, "NULL argument");
    const int wx = ntx - 2 * pml, wy = nty - 2 * pml, wz = ntz - 2 * pml;
    if (pml < 0 || wx < 1 || wy < 1 || wz < 1) return qvti_fail(QVTI_E_ARG, "grid too small for pml %d", pml);
    auto clampw = [](int i, int lo, int w) { return std::min(std::max(i - lo, 0), w - 1); };
    parallel_for(ntx, [&](long long ix) {
        const long long sx = clampw((int)ix, pml, wx);
        for (int iy = 0; iy < nty; ++iy) {
            const long long sy = clampw(iy, pml, wy);
            const float* src = interior + (sx * wy + sy) * wz;
            float* dst = total + (ix * nty + iy) * (long long)ntz;
            for (int iz = 0; iz < ntz; ++iz) dst[iz] = src[clampw(iz, pml, wz)];
        }
    });
    return QVTI_OK;
}

// Read an interior model (raw FP32, ix slowest, (ntx-2pml)(nty-2pml)(ntz-2pml)
// values - the reference's input files, viscoelas_3D.cpp:541-546) and pad it
// into the total grid.
extern "C" int qvti_load_model_raw(const char* path, int ntx, int nty, int ntz, int pml, float* total)
{
    if (!path || !total) return qvti_fail(QVTI_E_ARG, "NULL argument");
    const long long wx = ntx - 2LL * pml, wy = nty - 2LL * pml, wz = ntz - 2LL * pml;
    if (pml < 0 || wx < 1 || wy < 1 || wz < 1) return qvti_fail(QVTI_E_ARG, "grid too small for pml %d", pml);
    const long long n = wx * wy * wz;
    std::vector<float> in(n);
    FILE* f = fopen(path, "rb");
    if (!f) return qvti_fail(QVTI_E_ARG, "cannot open %s", path);
    const size_t got = fread(in.data(), sizeof(float), (size_t)n, f);
    fclose(f);
    if ((long long)got != n)
        return qvti_fail(QVTI_E_ARG, "%s holds %zu floats, the interior needs %lld", path, got, n);
    return qvti_pad_model(in.data(), ntx, nty, ntz, pml, total);
}

// Write n FP32 values raw: record gathers [nt][R] as the reference dumps them
// (viscoelas_3D.cpp:416-434), or interior snapshots [window] (ix slowest,
// cuviscoelas_3D.cu:1615-1627).
extern "C" int qvti_write_raw(const char* path, const float* data, long long n)
{
    if (!path || (!data && n > 0) || n < 0) return qvti_fail(QVTI_E_ARG, "bad arguments");
    FILE* f = fopen(path, "wb");
    if (!f) return qvti_fail(QVTI_E_ARG, "cannot create %s", path);
    const size_t put = n ? fwrite(data, sizeof(float), (size_t)n, f) : 0;
    const int closed = fclose(f);
    if ((long long)put != n || closed != 0) return qvti_fail(QVTI_E_ARG, "short write to %s", path);
    return QVTI_OK;
}
