// qvti_shots.cpp -- C++ host of the engine: the reference host's forward
// modelling driver (viscoelas_3D.cpp:26-480) for one process and one GPU,
// calling the C ABI (include/qvti.h) where the reference calls its CUDA
// entry points.
//
// Same steps, same conventions as the reference's main():
//   * grid: interior nx, ny, nz plus a pml-cell C-PML on every face; the
//     engine runs on the total grid ntx = nx + 2 pml, ... (viscoelas_3D.cpp:115-124);
//   * wavelet: Ricker of f0 (getrik, :494-502); w0 = 2 pi f0 (:109);
//   * shots: an nsx x nsy grid, dsx/dsy apart, starting nsx0/nsy0 cells
//     from the interior's corner, at depth `depth` below the C-PML (:152-186);
//   * receivers: every dr-th interior cell, rnx_max = nx/dr by
//     rny_max = ny/dr, r_ix = pml + ir/rny_max*dr, r_iy = pml + ir%rny_max*dr (:181-185);
//   * model: vp from a raw interior file padded by edge replication
//     (get_velQ, :537-655), or a constant; vs = vp/1.5, Qp = 200, Qs = 150,
//     rho = 1800 as the reference derives them (:664-673); gamma =
//     atan(1/Q)/pi (:285-291);
//   * per shot: qvti_forward_acoustic_3D (cuda_forward_acoustic_3D, :398-405),
//     then the three records written as raw FP32 [nt][R] (:416-434).
// Shots are stepped one after another on the chosen GPU; the reference's
// MPI shot partitioning (:236-256) becomes one process per GPU with
// --shot-begin/--shot-step.
//
// --rtm-type T migrates every shot after modelling it: the mode the
// reference declares with RTM_type, imaging_condition, direcr_or_not and
// check_steps (:42-65, 223-229) and never builds.  T = 1 Q-compensated
// (both wavefields propagated with the loss turned into gain), 2 with the
// viscoelastic scheme, 3 lossless; --imaging-condition 1 EXA, 3 source-
// normalised, 4 cross-correlation (2, partial correlation, is not built);
// --direct-vp V removes the direct wave by subtracting a run in a
// homogeneous vp = V model (a VTI model keeps its eps and delta); --mig-vp FILE migrates in another interior vp
// model (default: the modelling one).  The image and illumination of the
// total grid are written as raw FP32 [ntx][nty][ntz].
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "qvti.h"

namespace {

struct Options {
    int nx = 0, ny = 0, nz = 0, pml = 20, nt = 1000;
    float dx = 20.f, dy = 20.f, dz = 20.f, dt = 8e-4f, f0 = 20.f;
    int nsx = 1, nsy = 1, dsx = 30, dsy = 30, nsx0 = -1, nsy0 = -1, depth = 1, dr = 1;
    std::string vp_file;
    float vp_const = 2500.f;
    float vs_ratio = 1.5f, qp = 200.f, qs = 150.f, rho = 1800.f;
    std::string eps_file, delta_file;  // VTI (an extension: the reference is isotropic)
    float kspace_vref = 0.f;
    int media = 1;  // the reference's media_type: 1 viscoelastic, 2 dispersion-only, 3 lossless, 4 Q-compensating
    std::string out = ".";
    int device = 0, shot_begin = 0, shot_step = 1;
    // migration (0: modelling only)
    int snap_every = 0, snap_field = 2;  // snapshots of the interior (the reference's disabled dump)
    int rtm_type = 0, imaging_condition = 4, checkpoint = 500;
    float direct_vp = 0.f, rtm_eps = 0.01f;
    std::string mig_vp_file;
};

void usage(const char* prog)
{
    std::printf(
        "usage: %s --nx N --ny N --nz N [options]\n"
        "  grid    --pml 20 --nt 1000 --dx 20 --dy 20 --dz 20 --dt 8e-4 --f0 20\n"
        "  shots   --nsx 1 --nsy 1 --dsx 30 --dsy 30 --nsx0 nx/2 --nsy0 ny/2 --depth 1 --dr 1\n"
        "  model   --vp FILE (interior, raw FP32, ix slowest) | --vp-const 2500\n"
        "          --vs-ratio 1.5 --qp 200 --qs 150 --rho 1800 [--eps FILE --delta FILE (VTI)]\n"
        "  scheme  --kspace-vref V (opt-in k-space correction; 0 = the reference scheme)\n"
        "          --media 1 (1 viscoelastic, 2 dispersion-only, 3 lossless, 4 Q-compensating)\n"
        "  run     --out DIR --device 0 --shot-begin 0 --shot-step 1\n"
        "          --snap-every 0 (K > 0: interior snapshots every K steps) --snap-field 2 (0-8: vx..pyz)\n"
        "  rtm     --rtm-type 0|1|2|3 (0: modelling only) --imaging-condition 4 --checkpoint 500\n"
        "          --rtm-eps 0.01 --direct-vp 0 (homogeneous vp for direct-wave removal) --mig-vp FILE\n"
        "Writes DIR/vx_shot<is>.bin, vy_..., vz_... (raw FP32 [nt][rnx_max*rny_max]) and\n"
        "DIR/source_xy.txt; with --snap-every, DIR/snap_<field>_shot<is>.bin ([nsnap][nx][ny][nz]);\n"
        "with --rtm-type, DIR/image_shot<is>.bin, illum_shot<is>.bin and\n"
        "image_stack.bin (raw FP32 [ntx][nty][ntz]).  The total grid nx+2*pml, ... must be a\n"
        "supported FFT length.\n",
        prog);
}

bool parse(int argc, char** argv, Options& o)
{
    for (int i = 1; i < argc; ++i) {
        const std::string k = argv[i];
        if (k == "--help" || k == "-h") return false;
        if (i + 1 >= argc) {
            std::fprintf(stderr, "missing value for %s\n", k.c_str());
            return false;
        }
        const char* v = argv[++i];
        auto I = [&](int& x) { x = std::atoi(v); };
        auto F = [&](float& x) { x = (float)std::atof(v); };
        if (k == "--nx") I(o.nx);
        else if (k == "--ny") I(o.ny);
        else if (k == "--nz") I(o.nz);
        else if (k == "--pml") I(o.pml);
        else if (k == "--nt") I(o.nt);
        else if (k == "--dx") F(o.dx);
        else if (k == "--dy") F(o.dy);
        else if (k == "--dz") F(o.dz);
        else if (k == "--dt") F(o.dt);
        else if (k == "--f0") F(o.f0);
        else if (k == "--nsx") I(o.nsx);
        else if (k == "--nsy") I(o.nsy);
        else if (k == "--dsx") I(o.dsx);
        else if (k == "--dsy") I(o.dsy);
        else if (k == "--nsx0") I(o.nsx0);
        else if (k == "--nsy0") I(o.nsy0);
        else if (k == "--depth") I(o.depth);
        else if (k == "--dr") I(o.dr);
        else if (k == "--vp") o.vp_file = v;
        else if (k == "--vp-const") F(o.vp_const);
        else if (k == "--vs-ratio") F(o.vs_ratio);
        else if (k == "--qp") F(o.qp);
        else if (k == "--qs") F(o.qs);
        else if (k == "--rho") F(o.rho);
        else if (k == "--eps") o.eps_file = v;
        else if (k == "--delta") o.delta_file = v;
        else if (k == "--kspace-vref") F(o.kspace_vref);
        else if (k == "--media") I(o.media);
        else if (k == "--out") o.out = v;
        else if (k == "--device") I(o.device);
        else if (k == "--shot-begin") I(o.shot_begin);
        else if (k == "--shot-step") I(o.shot_step);
        else if (k == "--snap-every") I(o.snap_every);
        else if (k == "--snap-field") I(o.snap_field);
        else if (k == "--rtm-type") I(o.rtm_type);
        else if (k == "--imaging-condition") I(o.imaging_condition);
        else if (k == "--checkpoint") I(o.checkpoint);
        else if (k == "--rtm-eps") F(o.rtm_eps);
        else if (k == "--direct-vp") F(o.direct_vp);
        else if (k == "--mig-vp") o.mig_vp_file = v;
        else {
            std::fprintf(stderr, "unknown option %s\n", k.c_str());
            return false;
        }
    }
    if (o.nx < 1 || o.ny < 1 || o.nz < 1 || o.nt < 1 || o.dr < 1 || o.shot_step < 1) {
        std::fprintf(stderr, "--nx, --ny, --nz, --nt and --dr must be positive\n");
        return false;
    }
    if (o.media < 1 || o.media > 4) {
        std::fprintf(stderr, "--media must be 1, 2, 3 or 4\n");
        return false;
    }
    if (o.snap_every < 0 || o.snap_field < 0 || o.snap_field > 8) {
        std::fprintf(stderr, "--snap-every must be non-negative and --snap-field in 0..8\n");
        return false;
    }
    if (o.rtm_type < 0 || o.rtm_type > 3) {
        std::fprintf(stderr, "--rtm-type must be 0 (modelling only), 1, 2 or 3\n");
        return false;
    }
    if (o.imaging_condition != 1 && o.imaging_condition != 3 && o.imaging_condition != 4) {
        std::fprintf(stderr, "--imaging-condition must be 1 (EXA), 3 (source-normalised) or 4 (cross-correlation);"
                             " 2 (partial correlation) is not built\n");
        return false;
    }
    if (o.checkpoint < 1 || !(o.rtm_eps >= 0.f) || !(o.direct_vp >= 0.f)) {
        std::fprintf(stderr, "--checkpoint must be positive, --rtm-eps and --direct-vp non-negative\n");
        return false;
    }
    if (o.eps_file.empty() != o.delta_file.empty()) {
        std::fprintf(stderr, "--eps and --delta go together\n");
        return false;
    }
    if (o.nsx0 < 0) o.nsx0 = o.nx / 2;
    if (o.nsy0 < 0) o.nsy0 = o.ny / 2;
    return true;
}

int fail(const char* what)
{
    std::fprintf(stderr, "qvti_shots: %s: %s\n", what, qvti_last_error());
    return 1;
}

}  // namespace

int main(int argc, char** argv)
{
    Options o;
    if (!parse(argc, argv, o)) {
        usage(argv[0]);
        return argc > 1 && std::strcmp(argv[1], "--help") != 0 && std::strcmp(argv[1], "-h") != 0 ? 2 : 0;
    }
    const int pml = o.pml;
    const int ntx = o.nx + 2 * pml, nty = o.ny + 2 * pml, ntz = o.nz + 2 * pml;
    const long long ntp = (long long)ntx * nty * ntz;
    const float pi = 3.1415926f;  // viscoelas_3D.cpp:14
    const float w0 = 2 * pi * o.f0;
    const bool vti = !o.eps_file.empty();
    std::printf("ntx=%d, nty=%d, ntz=%d\nnx=%d, ny=%d, nz=%d, nt=%d, ns=%d\n", ntx, nty, ntz, o.nx, o.ny, o.nz,
                o.nt, o.nsx * o.nsy);

    // wavelet
    std::vector<float> rik(o.nt);
    if (qvti_ricker(o.nt, o.dt, o.f0, rik.data())) return fail("wavelet");

    // model on the total grid
    std::vector<float> velp(ntp), vels(ntp), Qp(ntp), Qs(ntp), rho(ntp), gp(ntp), gs(ntp), eps, del;
    if (!o.vp_file.empty()) {
        if (qvti_load_model_raw(o.vp_file.c_str(), ntx, nty, ntz, pml, velp.data())) return fail("vp model");
    } else {
        for (auto& v : velp) v = o.vp_const;
    }
    for (long long ip = 0; ip < ntp; ++ip) {
        vels[ip] = velp[ip] / o.vs_ratio;
        Qp[ip] = o.qp;
        Qs[ip] = o.qs;
        rho[ip] = o.rho;
    }
    if (qvti_gamma_from_q(Qp.data(), gp.data(), ntp) || qvti_gamma_from_q(Qs.data(), gs.data(), ntp))
        return fail("gamma");
    if (vti) {
        eps.resize(ntp);
        del.resize(ntp);
        if (qvti_load_model_raw(o.eps_file.c_str(), ntx, nty, ntz, pml, eps.data())) return fail("eps model");
        if (qvti_load_model_raw(o.delta_file.c_str(), ntx, nty, ntz, pml, del.data())) return fail("delta model");
    }
    float velp_max = 0.f;
    for (float v : velp) velp_max = v > velp_max ? v : velp_max;

    // observation system
    const int ns = o.nsx * o.nsy;
    const int rnx_max = o.nx / o.dr, rny_max = o.ny / o.dr;
    const int rnmax = rnx_max * rny_max;
    std::vector<int> r_ix(rnmax), r_iy(rnmax);
    if (qvti_reference_receivers(pml, rnx_max, rny_max, o.dr, r_ix.data(), r_iy.data())) return fail("receivers");
    std::vector<qvti_ref_source> ss(ns);
    for (int isx = 0; isx < o.nsx; ++isx)
        for (int isy = 0; isy < o.nsy; ++isy) {
            qvti_ref_source& s = ss[isx * o.nsy + isy];
            s.s_ix = isx * o.dsx + pml + o.nsx0;
            s.s_iy = isy * o.dsy + pml + o.nsy0;
            s.s_iz = pml + o.depth;
            s.r_iz = pml + o.depth;
            s.r_n = rnmax;
            s.r_ix = r_ix.data();
            s.r_iy = r_iy.data();
        }
    if (FILE* f = std::fopen((o.out + "/source_xy.txt").c_str(), "w")) {
        for (const auto& s : ss) std::fprintf(f, "%10d %10d\n", s.s_ix - pml, s.s_iy - pml);
        std::fclose(f);
    }

    // engine context (replaces getdevice / cuda_Device_malloc / cuda_Host_initialization)
    int ngpu = 0;
    if (qvti_device_count(&ngpu) || ngpu < 1) return fail("device count");
    qvti_grid g{ntx, nty, ntz, pml, o.dx, o.dy, o.dz, o.dt, o.f0, w0, vti ? 1 : 0};
    qvti_ctx* ctx = nullptr;
    if (qvti_create(&ctx, &g, 1, o.device % ngpu)) return fail("context");
    int rc = 0;
    // media_type (viscoelas_3D.cpp:57-60) -> QVTI_MEDIA_*: the modelling medium
    if (qvti_set_media(ctx, o.media - 1)) rc = fail("medium");
    if (!rc && vti &&
        qvti_set_model(ctx, velp.data(), gp.data(), vels.data(), gs.data(), rho.data(), eps.data(), del.data()))
        rc = fail("VTI model");
    if (!rc && o.kspace_vref > 0.f && qvti_set_kspace(ctx, o.kspace_vref, 0.f, 0.f, 0.f)) rc = fail("k-space mode");

    // migration contexts (source and receiver wavefields) on the migration
    // model, and the direct-wave context on a homogeneous one
    qvti_ctx *msrc = nullptr, *mrcv = nullptr, *dctx = nullptr;
    std::vector<float> stack;
    auto model_on = [&](qvti_ctx* c, int media, const std::vector<float>& vp) -> int {
        std::vector<float> vs(ntp);
        for (long long ip = 0; ip < ntp; ++ip) vs[ip] = vp[ip] / o.vs_ratio;
        if (qvti_set_media(c, media)) return fail("medium");
        if (qvti_set_model(c, vp.data(), gp.data(), vs.data(), gs.data(), rho.data(), vti ? eps.data() : nullptr,
                           vti ? del.data() : nullptr))
            return fail("model");
        if (o.kspace_vref > 0.f && qvti_set_kspace(c, o.kspace_vref, 0.f, 0.f, 0.f)) return fail("k-space mode");
        return 0;
    };
    if (!rc && o.rtm_type > 0) {
        // RTM_type 1: Q-compensated; 2: viscoelastic; 3: lossless (viscoelas_3D.cpp:42-45)
        const int media = o.rtm_type == 1 ? QVTI_MEDIA_QCOMP : (o.rtm_type == 2 ? QVTI_MEDIA_VISCO : QVTI_MEDIA_LOSSLESS);
        std::vector<float> mvp = velp;
        if (!o.mig_vp_file.empty() && qvti_load_model_raw(o.mig_vp_file.c_str(), ntx, nty, ntz, pml, mvp.data()))
            rc = fail("migration vp model");
        if (!rc && (qvti_create(&msrc, &g, 1, o.device % ngpu) || qvti_create(&mrcv, &g, 1, o.device % ngpu)))
            rc = fail("migration contexts");
        if (!rc) rc = model_on(msrc, media, mvp);
        if (!rc) rc = model_on(mrcv, media, mvp);
        if (!rc && o.direct_vp > 0.f) {
            if (qvti_create(&dctx, &g, 1, o.device % ngpu)) rc = fail("direct-wave context");
            if (!rc) rc = model_on(dctx, QVTI_MEDIA_VISCO, std::vector<float>(ntp, o.direct_vp));
        }
        stack.assign(ntp, 0.f);
    }

    // snapshots of the interior window every snap_every steps (the reference's
    // disabled dump, cuviscoelas_3D.cu:1615-1627), written per shot
    std::vector<float> snaps;
    if (!rc && o.snap_every > 0) {
        const long long nsnap = (o.nt + o.snap_every - 1) / o.snap_every;
        snaps.resize((size_t)(nsnap * o.nx * o.ny * o.nz));
        if (qvti_set_snapshot(ctx, o.snap_field, o.snap_every, snaps.data(), (long long)snaps.size()))
            rc = fail("snapshots");
    }

    // forward modelling, shot by shot (the entry point takes the whole source
    // array and the shot index, as cuda_forward_acoustic_3D does)
    std::vector<float> rec1((size_t)o.nt * rnmax), rec2(rec1.size()), rec3(rec1.size());
    for (int is = o.shot_begin; is < ns && !rc; is += o.shot_step) {
        const auto t0 = std::chrono::steady_clock::now();
        if (vti) {
            // the reference-compatible entry point takes the isotropic model;
            // a VTI model goes through qvti_forward with the same geometry
            qvti_shot sh{QVTI_SRC_FX, ss[is].s_ix, ss[is].s_iy, ss[is].s_iz + 10, rnmax, r_ix.data(), r_iy.data(),
                         nullptr};
            std::vector<int> rz(rnmax, ss[is].r_iz);
            sh.rz = rz.data();
            if (qvti_forward(ctx, o.nt, rik.data(), &sh, rec1.data(), rec2.data(), rec3.data())) rc = fail("forward");
        } else if (qvti_forward_acoustic_3D(ctx, 0, is, o.nt, ntx, nty, ntz, (int)ntp, o.nx, o.ny, o.nz, pml, o.dx,
                                            o.dy, o.dz, o.dt, o.f0, w0, velp_max, rik.data(), velp.data(), gp.data(),
                                            vels.data(), gs.data(), rho.data(), ss.data(), rnmax, rnx_max, rny_max,
                                            o.dr, rec1.data(), rec2.data(), rec3.data())) {
            rc = fail("forward");
        }
        if (rc) break;
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        const char* names[3] = {"vx", "vy", "vz"};
        const float* recs[3] = {rec1.data(), rec2.data(), rec3.data()};
        for (int k = 0; k < 3 && !rc; ++k) {
            const std::string path = o.out + "/" + names[k] + "_shot" + std::to_string(is) + ".bin";
            if (qvti_write_raw(path.c_str(), recs[k], (long long)o.nt * rnmax)) rc = fail("record");
        }
        if (!rc && !snaps.empty()) {
            static const char* fields[9] = {"vx", "vy", "vz", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz"};
            const std::string path =
                o.out + "/snap_" + fields[o.snap_field] + "_shot" + std::to_string(is) + ".bin";
            if (qvti_write_raw(path.c_str(), snaps.data(), (long long)snaps.size())) rc = fail("snapshot file");
        }
        std::printf("shot %d: %.3f s, %.3f Gcell-steps/s\n", is, s, (double)ntp * o.nt / s / 1e9);
        std::fflush(stdout);
        if (!rc && o.rtm_type > 0) {
            // the shot as the modelling ran it: x-force at s_iz + 10, receivers at r_iz
            std::vector<int> rz(rnmax, ss[is].r_iz);
            const qvti_shot sh{QVTI_SRC_FX, ss[is].s_ix, ss[is].s_iy, ss[is].s_iz + 10, rnmax, r_ix.data(),
                               r_iy.data(), rz.data()};
            if (dctx) {
                std::vector<float> d1(rec1.size()), d2(rec1.size()), d3(rec1.size());
                if (qvti_forward(dctx, o.nt, rik.data(), &sh, d1.data(), d2.data(), d3.data())) {
                    rc = fail("direct wave");
                    break;
                }
                for (size_t k = 0; k < rec1.size(); ++k) {
                    rec1[k] -= d1[k];
                    rec2[k] -= d2[k];
                    rec3[k] -= d3[k];
                }
            }
            const auto t1 = std::chrono::steady_clock::now();
            const qvti_rtm_opts ro{o.imaging_condition, QVTI_FIELD_PRESSURE, o.checkpoint, o.rtm_eps};
            std::vector<float> image(ntp), illum(ntp);
            if (qvti_rtm(msrc, mrcv, o.nt, rik.data(), &sh, rec1.data(), rec2.data(), rec3.data(), &ro, image.data(),
                         illum.data())) {
                rc = fail("migration");
                break;
            }
            for (long long ip = 0; ip < ntp; ++ip) stack[ip] += image[ip];
            const std::string tag = "_shot" + std::to_string(is) + ".bin";
            if (qvti_write_raw((o.out + "/image" + tag).c_str(), image.data(), ntp) ||
                qvti_write_raw((o.out + "/illum" + tag).c_str(), illum.data(), ntp))
                rc = fail("image");
            std::printf("shot %d migrated: %.3f s\n", is,
                        std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count());
            std::fflush(stdout);
        }
    }
    if (!rc && o.rtm_type > 0 && qvti_write_raw((o.out + "/image_stack.bin").c_str(), stack.data(), ntp))
        rc = fail("image stack");
    if (dctx) qvti_destroy(dctx);
    if (mrcv) qvti_destroy(mrcv);
    if (msrc) qvti_destroy(msrc);
    qvti_destroy(ctx);  // replaces cuda_Device_free
    return rc;
}
