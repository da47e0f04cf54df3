// pfast_t3d -- 3D first-arrival traveltime tomography driver (the reference's
// T_PFAST3D: src/T_PFAST3D.cpp), same parameter file, input files, outputs,
// resume convention and iteration logic, with every gradient evaluation one
// batched GPU call over all sources (pfast_shots3d_*: FP32 sweeps, FP64 gradient
// and misfit) and the gradient smoothing on the GPU (pfast_regu3d).
//
// usage: pfast_t3d input.inp [LM=5] [--forward] [--max-iter K] [--no-autotune]
//   --forward     forward modelling only (the reference's -DNoInverse build):
//                 writes <prefix>_Tr3D.bin
//   --max-iter K  stop after K iterations (default 100, the reference's cap)
//   --no-autotune keep the default sweep variant and schedule (an inversion otherwise
//                 first times them on its shot set: pfast_shots3d_autotune and
//                 pfast_shots3d_autotune_schedule; results are the same)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "driver_common.hpp"

using drv::check;

int main(int argc, char **argv)
{
    std::vector<const char *> pos;
    bool forward = false, autotune = true;
    int MAXIter = 100;
    for (int a = 1; a < argc; a++) {
        if (!strcmp(argv[a], "--forward")) forward = true;
        else if (!strcmp(argv[a], "--no-autotune")) autotune = false;
        else if (!strcmp(argv[a], "--max-iter") && a + 1 < argc) MAXIter = atoi(argv[++a]);
        else pos.push_back(argv[a]);
    }
    if (pos.empty() || pos.size() > 2) {
        printf(" usage: pfast_t3d input_file [L-BFGS memory, default 5] [--forward] [--max-iter K] [--no-autotune]\n");
        return 1;
    }
    const int LM = pos.size() == 2 ? atoi(pos[1]) : 5;
    printf(" * 3-D traveltime tomography on the GPU (libpfast)\n");
    if (!forward)
        printf("Only Applicable to Optimization Method: L-BFGS\n\t...%s Memory Number m=%d\n",
               pos.size() == 2 ? "Specified" : "Default", LM);

    // ---- parameter file (src/T_PFAST3D.cpp:113-136)
    const std::vector<std::string> inp = drv::read_inp(pos[0]);
    const std::string vp0file = drv::tok(inp, 0, 0), srcfile = drv::tok(inp, 1, 0);
    const int sNX = atoi(drv::tok(inp, 2, 0).c_str()), sNY = atoi(drv::tok(inp, 2, 1).c_str());
    const std::string recfile = drv::tok(inp, 3, 0);
    const int rNX = atoi(drv::tok(inp, 4, 0).c_str()), rNY = atoi(drv::tok(inp, 4, 1).c_str());
    const int NX = atoi(drv::tok(inp, 5, 0).c_str()), NY = atoi(drv::tok(inp, 5, 1).c_str()),
              NZ = atoi(drv::tok(inp, 5, 2).c_str());
    const double dx = atof(drv::tok(inp, 5, 3).c_str()), dy = atof(drv::tok(inp, 5, 4).c_str()),
                 dz = atof(drv::tok(inp, 5, 5).c_str());
    std::string vpfile;
    double nux = 0, nuy = 0, nuz = 0;
    int lsrc = 3;
    if (!forward) {
        vpfile = drv::tok(inp, 6, 0);
        nux = atof(drv::tok(inp, 7, 0).c_str());
        nuy = atof(drv::tok(inp, 7, 1).c_str());
        nuz = atof(drv::tok(inp, 7, 2).c_str());
        lsrc = atoi(drv::tok(inp, 8, 1).c_str());
    }
    // a model file name ending in a digit resumes after that iteration
    // (src/T_PFAST3D.cpp:138-145)
    int isresume = 0;
    printf("\nThe initial velocity file... %s\n", vp0file.c_str());
    if (!vp0file.empty() && vp0file.back() != 'p') {
        isresume = vp0file.back() - '0';
        printf("|==========This is a RESUME run===========|\n");
    } else {
        printf("|==========This is a NEW run===========|\n");
    }
    printf("The size of the model... %d X %d X %d\nSample Interval... %.2lfm x %.2lfm x %.2lfm\n",
           NX, NY, NZ, dx, dy, dz);

    // ---- inputs (src/T_PFAST3D.cpp:196-320)
    const long N = (long)NX * NY * NZ;
    std::vector<double> vp, srcArray, recArray;
    drv::readmod(vp, vp0file, N);
    drv::readmod(srcArray, srcfile, (size_t)sNX * sNY);
    drv::readmod(recArray, recfile, (size_t)rNX * rNY);
    const int NumofSrc = sNX, NumofRec = rNX;
    std::vector<int> src(3 * (size_t)NumofSrc), srnum(NumofSrc), rec(3 * (size_t)NumofRec);
    std::vector<double> nxyz(3 * (size_t)NumofRec), T_obs(NumofRec);
    for (int ii = 0; ii < NumofSrc; ii++) {
        const double *s = &srcArray[(size_t)ii * sNY];
        src[3 * ii + 0] = drv::map_index(s[0], dy);
        src[3 * ii + 1] = drv::map_index(s[1], dx);
        src[3 * ii + 2] = drv::map_index(s[2], dz);
        const int i = src[3 * ii], j = src[3 * ii + 1], k = src[3 * ii + 2];
        if (i < 0 || i >= NX || j < 0 || j >= NY || k < 0 || k >= NZ ||
            vp[(long)k * NX * NY + (long)i * NY + j] < 400) {
            printf("Source %d:\t%d\t%d\t%d is not in the model volume, please check...\n", ii, i + 1,
                   j + 1, k + 1);
            return 1;
        }
        srnum[ii] = (int)ceil(s[3]);
        if (ii <= 20 || ii == NumofSrc - 1)
            printf("| %d\t%d\t%d\t%d\t%d\n", ii, i + 1, j + 1, k + 1, srnum[ii]);
        else if (ii == 21)
            printf("| ......\n");
    }
    for (int ii = 0; ii < NumofRec; ii++) {
        const double *r = &recArray[(size_t)ii * rNY];
        rec[3 * ii + 0] = drv::map_index(r[0], dy);
        rec[3 * ii + 1] = drv::map_index(r[1], dx);
        rec[3 * ii + 2] = drv::map_index(r[2], dz);
        const int i = rec[3 * ii], j = rec[3 * ii + 1], k = rec[3 * ii + 2];
        if (i < 0 || i >= NX || j < 0 || j >= NY || k < 0 || k >= NZ ||
            vp[(long)k * NX * NY + (long)i * NY + j] < 400) {
            printf("Receiver %d:\t%d\t%d\t%d is not in the model volume, please check...\n", ii,
                   i + 1, j + 1, k + 1);
            return 1;
        }
        nxyz[3 * ii] = r[3];
        nxyz[3 * ii + 1] = r[4];
        nxyz[3 * ii + 2] = r[5];
        T_obs[ii] = rNY > 6 ? r[6] : 150.0;
    }
    long need = 0;
    for (int s = 0; s < NumofSrc; s++) need += srnum[s];
    if (need > NumofRec) {
        printf("the source table asks for %ld receiver rows, the receiver file has %d\n", need,
               NumofRec);
        return 1;
    }
    const std::string p_vpfile = drv::prefix(vp0file, isresume == 0 ? 3 : isresume <= 9 ? 4 : 5);

    pfast_ctx *ctx = nullptr;
    check(pfast_ctx_create(0, &ctx), "pfast_ctx_create");
    // a cycle's source chunks split over 3 streams: same results, more launches in flight
    check(pfast_ctx_set_streams(ctx, 3), "pfast_ctx_set_streams");
    // replay the sweep cycles between convergence checks from captured hipGraphs (same
    // results, about 1% less time per gradient evaluation)
    check(pfast_ctx_set_graphs(ctx, 1), "pfast_ctx_set_graphs");
    pfast_model3d *model = nullptr;
    check(pfast_model3d_create(ctx, vp.data(), NX, NY, NZ, dx, dy, dz, &model), "pfast_model3d_create");
    pfast_shots3d *shots = nullptr;
    check(pfast_shots3d_create(model, NumofSrc, src.data(), srnum.data(), rec.data(), nxyz.data(),
                               T_obs.data(), &shots),
          "pfast_shots3d_create");
    const int MAXSwp = 100;

    if (forward) {  // src/T_PFAST3D.cpp:658-689: one pick per receiver row, sources in order
        check(pfast_shots3d_run(shots, PFAST_FWD_ONLY, MAXSwp), "pfast_shots3d_run");
        std::vector<double> picks(need);
        check(pfast_shots3d_picks(shots, picks.data()), "pfast_shots3d_picks");
        drv::writemod(picks.data(), p_vpfile + "_Tr3D.bin", picks.size());
        pfast_shots3d_destroy(shots);
        pfast_model3d_destroy(model);
        pfast_ctx_destroy(ctx);
        return 0;
    }

    if (autotune) {  // every candidate gives the same results: the choice is by time
        int variant = 0, schedule = 0, interval = 0;
        // the schedule tuner keeps the setting in use unless another is 2% faster: start
        // it from a check every 2 cycles, the fastest on large shot sets
        check(pfast_ctx_set_check_interval(ctx, 2), "pfast_ctx_set_check_interval");
        check(pfast_shots3d_autotune(shots, nullptr, 0, &variant, nullptr), "pfast_shots3d_autotune");
        check(pfast_shots3d_autotune_schedule(shots, &schedule, &interval, nullptr),
              "pfast_shots3d_autotune_schedule");
        printf(" * autotune: sweep variant %d (%s), schedule %d, convergence check every %d "
               "cycle(s)\n", variant, pfast_sweep_name(variant), schedule, interval);
    }

    drv::Optimiser opt;
    opt.init(ctx, N, LM);
    opt.schemetag = 3;  // the reference fixes L-BFGS (src/T_PFAST3D.cpp:428)
    opt.lsrc = lsrc;
    opt.MAXLineSch = 2;
    opt.dim2 = false;
    std::vector<double> global_E;
    std::vector<double> g(N);
    double E = 0;
    auto evaluate = [&](const std::vector<double> &m) {
        check(pfast_model3d_set_vp(model, m.data()), "set_vp");
        check(pfast_shots3d_run(shots, PFAST_FWD_ADJ, MAXSwp), "pfast_shots3d_run");
        check(pfast_shots3d_gradient(shots, g.data(), &E), "pfast_shots3d_gradient");
    };
    drv::EvalFn trial = [&](const std::vector<double> &m, std::vector<double> &Dj, double &phi) {
        evaluate(m);
        Dj = g;
        phi = E;
        check(pfast_regu3d(Dj.data(), vp.data(), nux, nuy, nuz, dx, dy, dz, NX, NY, NZ), "regu3d");
    };
    auto save_alpha = [](double a) {
        FILE *fe = fopen("alpha1.tmp", "w");
        if (fe) {
            fprintf(fe, "%g\n", a);
            fclose(fe);
        }
    };

    int Iterk = 0;
    if (isresume) {  // src/T_PFAST3D.cpp:345-356
        drv::readmod(opt.Djk1, p_vpfile + "_bf_regu_" + std::to_string(isresume) + ".vp", N);
        opt.dk = opt.Djk1;
        Iterk += isresume;
        printf("Iterating No. %d resumed: (nux,nuy,nuz)=(%f,%f,%f)\n", Iterk, nux, nuy, nuz);
    }
    bool convflag = false;
    while (!convflag && Iterk++ < MAXIter) {
        printf("Iterating No. %d started: (nux,nuy,nuz)=(%f,%f,%f)\n", Iterk, nux, nuy, nuz);
        opt.Djk0 = opt.Djk1;
        const std::string k = std::to_string(Iterk);
        if (isresume) {  // src/T_PFAST3D.cpp:365-384
            drv::readmod(opt.Djk1, p_vpfile + "_bf_regu_" + k + ".vp", N);
            FILE *fe = fopen("global_E.tmp", "r");
            global_E.clear();
            if (fe) {
                char line[256];
                while (fgets(line, sizeof line, fe)) global_E.push_back(atof(line));
                fclose(fe);
            } else {
                printf(" The file global_E.tmp cannot be opened\n");
            }
            fe = fopen("alpha1.tmp", "r");
            if (fe) {
                char line[256];
                while (fgets(line, sizeof line, fe)) opt.alpha1 = atof(line);
                fclose(fe);
            } else {
                printf(" The file alpha1.tmp cannot be opened\n");
            }
            global_E.resize(Iterk, 0.0);
            printf("Average Time Residual: %g ms\n", 1000 * sqrt(2 * global_E[Iterk - 1] / NumofRec));
            isresume = 0;
        } else {
            evaluate(vp);
            opt.Djk1 = g;
            global_E.resize(Iterk - 1);
            global_E.push_back(E);
            printf("Average Time Residual: %g ms\n", 1000 * sqrt(2 * E / NumofRec));
            drv::writemod(opt.Djk1.data(), p_vpfile + "_bf_regu_" + k + ".vp", N);
            check(pfast_regu3d(opt.Djk1.data(), vp.data(), nux, nuy, nuz, dx, dy, dz, NX, NY, NZ),
                  "regu3d");
            if (Iterk == 1) opt.alpha1 = 5e2 * pfast_magnitude_of(opt.Djk1.data(), N);
            drv::writemod(opt.Djk1.data(), p_vpfile + "_af_regu_" + k + ".vp", N);
            FILE *fe = fopen((p_vpfile + "_GlobalErr.txt").c_str(), "w");
            if (fe) {
                for (int i = 0; i < Iterk; i++) fprintf(fe, "%g\n", global_E[i]);
                fclose(fe);
            }
        }
        opt.direction(Iterk);
        opt.line_search(vp, global_E[Iterk - 1], trial, save_alpha);
        drv::updatevp(vp, opt.dk, opt.alpha1, 1400, 8000);
        drv::writemod(vp.data(), p_vpfile + ".vp" + k, N);
        convflag = drv::check_conv(opt.dk, opt.alpha1, vp, global_E, Iterk, 1e-2, N);
        printf("--------Interating No. %d done---------------\n\n", Iterk);
    }
    drv::writemod(vp.data(), vpfile, N);
    FILE *fe = fopen((p_vpfile + "_AveRes_T.txt").c_str(), "w");
    if (fe) {
        for (size_t i = 0; i < global_E.size(); i++) fprintf(fe, "%g\n", 1000 * sqrt(2 * global_E[i] / NumofRec));
        fclose(fe);
    }
    pfast_shots3d_destroy(shots);
    pfast_model3d_destroy(model);
    pfast_ctx_destroy(ctx);
    return 0;
}
