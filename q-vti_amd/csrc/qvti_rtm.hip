// qvti_rtm.hip -- reverse-time migration on the engine.
//
// The reference declares an RTM mode it never builds: RTM_type (Q-compensated
// / viscoelastic data / lossless), imaging_condition (excitation amplitude,
// partial correlation, source-normalised and plain cross-correlation),
// media_type and wavefield reconstruction flags (viscoelas_3D.cpp:42-65) and
// checkpoint bookkeeping every check_steps = 500 steps (viscoelas_3D.cpp:
// 223-229).  This file builds it on the forward engine:
//
//  * source wavefield S(t), t = 0..nt-1: the shot's forward run on the source
//    context, checkpointed every C steps (qvti_checkpoint) and recomputed one
//    segment [kC, (k+1)C) at a time in reverse segment order, the imaged field
//    of every step kept in a segment buffer;
//  * receiver wavefield R(tau), tau = 0..nt-1: a forward run on the receiver
//    context (same grid; its model and medium - qvti_set_media, e.g.
//    Q-compensating - may differ) whose only sources are the records injected
//    time-reversed as forces at the receiver cells: after the stress update
//    of step tau, v_c[receiver r] += rec_c[nt-1-tau][r];
//  * imaging at equal physical time t = nt-1-tau:
//      cross-correlation      I = sum_t S(t) R(t)
//      source-normalised      I = sum_t S R / (sum_t S^2 + eps max sum_t S^2)
//      excitation amplitude   I = R(t_ex) / S(t_ex), t_ex the time of max |S|,
//                             where |S(t_ex)| > eps max |S(t_ex)|
//    (the partial-correlation condition is not built).
// The two contexts run on their own streams: segment k-1 is recomputed on the
// source stream while the receiver stream images segment k (two segment
// buffers, ordered by events).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "qvti_ctx.h"
#include "qvti_internal.h"

namespace {

#define RTMCHK(x)                                                                                       \
    do {                                                                                                \
        hipError_t e_ = (x);                                                                            \
        if (e_ != hipSuccess) {                                                                         \
            (void)hipGetLastError();                                                                    \
            rc = qvti_fail(e_ == hipErrorOutOfMemory ? QVTI_E_NOMEM : QVTI_E_DEVICE, "%s failed: %s", #x, \
                           hipGetErrorString(e_));                                                      \
            goto done;                                                                                  \
        }                                                                                               \
    } while (0)

// The imaged field of shot 0: one wavefield or the mean normal stress.
struct FieldRef {
    const float *a, *b, *c;  // b, c non-null: (a + b + c) / 3
};
FieldRef field_of(qvti_ctx* c, int field)
{
    if (field == QVTI_FIELD_PRESSURE) return {fld(c->sig, 0, c), fld(c->sig, 1, c), fld(c->sig, 2, c)};
    if (field < 3) return {fld(c->v, field, c), nullptr, nullptr};
    return {fld(c->sig, field - 3, c), nullptr, nullptr};
}

__device__ __forceinline__ float field_at(const FieldRef& f, long long i)
{
    return f.b ? (f.a[i] + f.b[i] + f.c[i]) / 3.0f : f.a[i];
}

__global__ __launch_bounds__(256) void k_field_copy(float* dst, FieldRef f, long long N)
{
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < N; i += (long long)gridDim.x * 256)
        dst[i] = field_at(f, i);
}

// Extra point sources: clear the previous step's entries of the dense
// injection buffer, then add this step's traces to v and to the buffer (the
// next velocity step adds the buffer to dv, as it does for the shot's source).
struct InjArgs {
    float* v[3];
    float* inj[3];
    const long long* ip;
    const float* tr[3];  // this step's row of each component's records (null: not injected)
    int R;
};
__global__ __launch_bounds__(256) void k_inj_clear(InjArgs a)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= a.R) return;
    for (int c = 0; c < 3; ++c)
        if (a.tr[c]) a.inj[c][a.ip[r]] = 0.f;
}
__global__ __launch_bounds__(256) void k_inj_add(InjArgs a)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= a.R) return;
    for (int c = 0; c < 3; ++c)
        if (a.tr[c]) {
            const float s = a.tr[c][r];
            atomicAdd(a.v[c] + a.ip[r], s);
            atomicAdd(a.inj[c] + a.ip[r], s);
        }
}

// cross-correlation (and source illumination) of one time level
__global__ __launch_bounds__(256) void k_img_xcorr(float* img, float* ill, const float* s, FieldRef r, long long N)
{
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < N; i += (long long)gridDim.x * 256) {
        const float sv = s[i];
        img[i] += sv * field_at(r, i);
        ill[i] += sv * sv;
    }
}

// excitation amplitude: track max |S| and its time step (first of equal maxima)
__global__ __launch_bounds__(256) void k_exa_track(float* smax, int* tex, FieldRef s, int t, long long N)
{
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < N; i += (long long)gridDim.x * 256) {
        const float v = field_at(s, i);
        if (fabsf(v) > fabsf(smax[i])) {
            smax[i] = v;
            tex[i] = t;
        }
    }
}
// (cells whose excitation amplitude is at most thr are not imaged: the
// division would amplify noise where the source wavefield never arrived)
__global__ __launch_bounds__(256) void k_img_exa(float* img, const float* smax, const int* tex, FieldRef r, int t,
                                                 float thr, long long N)
{
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < N; i += (long long)gridDim.x * 256)
        if (tex[i] == t && fabsf(smax[i]) > thr) img[i] += field_at(r, i) / smax[i];
}

unsigned grid_for(long long N) { return (unsigned)std::min<long long>((N + 255) / 256, 65535LL * 4); }

}  // namespace

extern "C" int qvti_rtm(qvti_ctx* cs, qvti_ctx* cr, int nt, const float* rik, const qvti_shot* shot,
                        const float* rec_vx, const float* rec_vy, const float* rec_vz, const qvti_rtm_opts* o,
                        float* image, float* illum)
{
    if (!cs || !cr || !rik || !shot || !o || !image) return qvti_fail(QVTI_E_ARG, "NULL argument");
    if (cs == cr) return qvti_fail(QVTI_E_ARG, "the source and receiver wavefields need two contexts");
    if (cs->g.nx != cr->g.nx || cs->g.ny != cr->g.ny || cs->g.nz != cr->g.nz || cs->g.pml != cr->g.pml ||
        cs->g.dt != cr->g.dt || cs->device != cr->device)
        return qvti_fail(QVTI_E_ARG, "source and receiver contexts must share grid, C-PML, dt and device");
    if (!cs->has_model || !cr->has_model) return qvti_fail(QVTI_E_STATE, "both contexts need a model");
    if (nt < 1) return qvti_fail(QVTI_E_ARG, "nt must be >= 1");
    if (o->condition != QVTI_IC_XCORR && o->condition != QVTI_IC_SRCNORM && o->condition != QVTI_IC_EXA)
        return qvti_fail(QVTI_E_ARG, "imaging condition %d not built (1 EXA, 3 source-normalised, 4 cross-correlation)",
                         o->condition);
    if (o->field < 0 || o->field > QVTI_FIELD_PRESSURE) return qvti_fail(QVTI_E_ARG, "field %d", o->field);
    if (o->checkpoint_every < 1) return qvti_fail(QVTI_E_ARG, "checkpoint_every must be >= 1");
    if (!(o->eps >= 0.f)) return qvti_fail(QVTI_E_ARG, "eps must be >= 0");
    const int R = shot->nrec;
    if (R < 1 || !shot->rx || !shot->ry || !shot->rz) return qvti_fail(QVTI_E_ARG, "the shot needs receivers");
    if (!rec_vx && !rec_vy && !rec_vz) return qvti_fail(QVTI_E_ARG, "no records to inject");
    const qvti_grid& g = cs->g;
    for (int r = 0; r < R; ++r)
        if (shot->rx[r] < 0 || shot->rx[r] >= g.nx || shot->ry[r] < 0 || shot->ry[r] >= g.ny || shot->rz[r] < 0 ||
            shot->rz[r] >= g.nz)
            return qvti_fail(QVTI_E_ARG, "receiver %d outside the grid", r);

    const long long N = cs->N;
    const int C = std::min(o->checkpoint_every, nt);
    const int K = (nt + C - 1) / C;  // segments
    const bool exa = o->condition == QVTI_IC_EXA;
    int rc = QVTI_OK;
    // runs without records (the wavefields only): the shot's receivers are the injection cells
    qvti_shot s0 = *shot;
    s0.nrec = 0;
    std::vector<float> silent(nt, 0.f);
    std::vector<std::vector<char>> ckpt(K);
    std::vector<long long> ip(R);
    for (int r = 0; r < R; ++r) ip[r] = ((long long)shot->rx[r] * g.ny + shot->ry[r]) * g.nz + shot->rz[r];
    float *d_inj = nullptr, *d_img = nullptr, *d_ill = nullptr, *d_seg = nullptr, *d_rec = nullptr, *d_smax = nullptr;
    int* d_tex = nullptr;
    long long* d_ip = nullptr;
    hipEvent_t ready[2] = {nullptr, nullptr}, used[2] = {nullptr, nullptr};
    const float* recs[3] = {rec_vx, rec_vy, rec_vz};
    const FieldRef fs = field_of(cs, o->field), fr = field_of(cr, o->field);
    const unsigned gN = grid_for(N), gR = (unsigned)((R + 255) / 256);
    float exa_thr = 0.f;

    RTMCHK(hipSetDevice(cs->device));
    RTMCHK(hipMalloc(&d_inj, sizeof(float) * 3 * cr->S * N));
    RTMCHK(hipMalloc(&d_img, sizeof(float) * N));
    RTMCHK(hipMalloc(&d_ill, sizeof(float) * N));
    RTMCHK(hipMalloc(&d_ip, sizeof(long long) * R));
    RTMCHK(hipMalloc(&d_rec, sizeof(float) * 3 * (size_t)nt * R));
    RTMCHK(hipMemset(d_inj, 0, sizeof(float) * 3 * cr->S * N));
    RTMCHK(hipMemset(d_img, 0, sizeof(float) * N));
    RTMCHK(hipMemset(d_ill, 0, sizeof(float) * N));
    RTMCHK(hipMemcpy(d_ip, ip.data(), sizeof(long long) * R, hipMemcpyHostToDevice));
    for (int c = 0; c < 3; ++c)
        if (recs[c]) RTMCHK(hipMemcpy(d_rec + (size_t)c * nt * R, recs[c], sizeof(float) * nt * R, hipMemcpyHostToDevice));
    if (exa) {
        RTMCHK(hipMalloc(&d_smax, sizeof(float) * N));
        RTMCHK(hipMalloc(&d_tex, sizeof(int) * N));
        RTMCHK(hipMemset(d_smax, 0, sizeof(float) * N));
        RTMCHK(hipMemset(d_tex, 0xff, sizeof(int) * N));  // -1: never excited
    } else {
        RTMCHK(hipMalloc(&d_seg, sizeof(float) * 2 * (size_t)C * N));
    }
    for (int b = 0; b < 2; ++b) {
        RTMCHK(hipEventCreateWithFlags(&ready[b], hipEventDisableTiming));
        RTMCHK(hipEventCreateWithFlags(&used[b], hipEventDisableTiming));
    }

    // ---- forward source pass: checkpoints at the segment starts (and the
    // excitation times for EXA)
    rc = qvti_begin(cs, nt, rik, &s0, 1);
    if (rc) goto done;
    for (int k = 0; k < K && !rc; ++k) {
        const int t0 = k * C, t1 = std::min(nt, t0 + C);
        if (k > 0) {
            const long long nb = qvti_checkpoint_bytes(cs);
            if (nb < 0) {
                rc = (int)nb;
                break;
            }
            ckpt[k].resize((size_t)nb);
            rc = qvti_checkpoint(cs, ckpt[k].data(), nb);
            if (rc) break;
        }
        if (exa) {
            for (int t = t0; t < t1 && !rc; ++t) {
                rc = qvti_step(cs, 1);
                hipLaunchKernelGGL(k_exa_track, dim3(gN), dim3(256), 0, cs->stream, d_smax, d_tex, fs, t, N);
            }
        } else {
            rc = qvti_step(cs, t1 - t0);
        }
    }
    if (!rc) rc = qvti_sync(cs);
    if (rc) goto done;
    if (exa) {
        // the excitation amplitudes below eps * max |S(t_ex)| are not imaged
        std::vector<float> sm(N);
        RTMCHK(hipMemcpy(sm.data(), d_smax, sizeof(float) * N, hipMemcpyDeviceToHost));
        float mx = 0.f;
        for (float v : sm) mx = std::max(mx, std::fabs(v));
        exa_thr = o->eps * mx;
    }

    // ---- receiver pass, segments in reverse order; the receiver context's
    // velocity kernel adds the injection buffer to dv
    for (int c = 0; c < 3; ++c) cr->a_vz.inj[c] = d_inj + (long long)c * cr->S * N;  // zero where not injected
    rc = qvti_begin(cr, nt, silent.data(), &s0, 1);
    if (rc) goto done;
    for (int k = K - 1; k >= 0 && !rc; --k) {
        const int t0 = k * C, t1 = std::min(nt, t0 + C), b = k & 1;
        float* seg = exa ? nullptr : d_seg + (size_t)b * C * N;
        if (!exa) {
            // recompute segment k of the source wavefield into buffer b once the
            // receiver stream has finished with what it held (segment k + 2)
            rc = qvti_begin(cs, nt, rik, &s0, 1);
            if (!rc && k > 0) rc = qvti_restore(cs, ckpt[k].data(), (long long)ckpt[k].size());
            if (rc) break;
            RTMCHK(hipStreamWaitEvent(cs->stream, used[b], 0));
            for (int t = t0; t < t1 && !rc; ++t) {
                rc = qvti_step(cs, 1);
                hipLaunchKernelGGL(k_field_copy, dim3(gN), dim3(256), 0, cs->stream, seg + (size_t)(t - t0) * N, fs, N);
            }
            if (rc) break;
            RTMCHK(hipEventRecord(ready[b], cs->stream));
            RTMCHK(hipStreamWaitEvent(cr->stream, ready[b], 0));
        }
        for (int t = t1 - 1; t >= t0 && !rc; --t) {
            rc = qvti_step(cr, 1);
            if (rc) break;
            InjArgs ia;
            for (int c = 0; c < 3; ++c) {
                ia.v[c] = fld(cr->v, c, cr);
                ia.inj[c] = d_inj + (long long)c * cr->S * N;
                ia.tr[c] = recs[c] ? d_rec + (size_t)c * nt * R + (size_t)t * R : nullptr;
            }
            ia.ip = d_ip;
            ia.R = R;
            hipLaunchKernelGGL(k_inj_clear, dim3(gR), dim3(256), 0, cr->stream, ia);
            hipLaunchKernelGGL(k_inj_add, dim3(gR), dim3(256), 0, cr->stream, ia);
            if (exa)
                hipLaunchKernelGGL(k_img_exa, dim3(gN), dim3(256), 0, cr->stream, d_img, d_smax, d_tex, fr, t,
                                   exa_thr, N);
            else
                hipLaunchKernelGGL(k_img_xcorr, dim3(gN), dim3(256), 0, cr->stream, d_img, d_ill,
                                   seg + (size_t)(t - t0) * N, fr, N);
        }
        if (!exa && !rc) RTMCHK(hipEventRecord(used[b], cr->stream));
    }
    if (!rc) rc = qvti_sync(cr);
    if (!rc) rc = qvti_sync(cs);
    if (rc) goto done;
    RTMCHK(hipGetLastError());
    {
        std::vector<float> ill(exa ? 0 : N);
        RTMCHK(hipMemcpy(image, d_img, sizeof(float) * N, hipMemcpyDeviceToHost));
        if (!exa) RTMCHK(hipMemcpy(ill.data(), d_ill, sizeof(float) * N, hipMemcpyDeviceToHost));
        if (o->condition == QVTI_IC_SRCNORM) {
            const float mx = *std::max_element(ill.begin(), ill.end());
            const float e = o->eps * mx;
            for (long long i = 0; i < N; ++i) image[i] = ill[i] + e > 0.f ? image[i] / (ill[i] + e) : 0.f;
        }
        if (illum && !exa) std::copy(ill.begin(), ill.end(), illum);
        if (illum && exa) RTMCHK(hipMemcpy(illum, d_smax, sizeof(float) * N, hipMemcpyDeviceToHost));
    }

done:
    if (cs->stream) hipStreamSynchronize(cs->stream);
    if (cr->stream) hipStreamSynchronize(cr->stream);
    for (int c = 0; c < 3; ++c) cr->a_vz.inj[c] = nullptr;
    cr->begun = 0;  // the graph captured with the injection buffer is stale
    cs->begun = 0;
    for (auto& e : ready)
        if (e) hipEventDestroy(e);
    for (auto& e : used)
        if (e) hipEventDestroy(e);
    hipFree(d_inj);
    hipFree(d_img);
    hipFree(d_ill);
    hipFree(d_seg);
    hipFree(d_rec);
    hipFree(d_smax);
    hipFree(d_tex);
    hipFree(d_ip);
    return rc;
}
