// pfast_rt2d -- 2D joint transmission / reflection traveltime tomography driver
// (the reference's RT_PFAST2D: src/RT_PFAST2D.cpp), same parameter file, input
// files, outputs and iteration logic, with every gradient evaluation run as one
// batched GPU call over all sources (pfast_shots2d_*) instead of the MPI per-rank
// source loops, and the gradient smoothing on the GPU (pfast_regu2d).
//
// usage: pfast_rt2d input.inp [LM=5 [RTtag=3]] [--forward] [--no-shot-gather]
//                   [--max-iter K] [--vmin V] [--vmax V]
//   --forward         forward modelling only (the reference's -DNoInverse build):
//                     writes <prefix>_Tr2D.bin and the per-source TTd / TTu fields
//   --no-shot-gather  with --forward, skip the per-source field files
//   --max-iter K      stop after K iterations (default 500, the reference's cap)
//   --vmin V, --vmax V  the velocity clamp of each model update (default the
//                     reference's MINV / MAXV = 1600 / 4500, src/RT_padjoint2d.h:17,21;
//                     a model slower than 1600 m/s is lifted to it by the first update)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "driver_common.hpp"

using drv::check;

namespace {

struct Survey2D {
    int NX = 0, NY = 0, NumofR = 0, NumofSrc = 0, NumofRec = 0;
    double dx = 0, dz = 0;
    std::vector<double> vp, Z, WR, nxy, TT_obs, TR_obs;
    std::vector<int> src, srnum, rec;
};

void usage()
{
    printf(" usage: pfast_rt2d input_file [Maximum Memory Number=5] [Tomography Scheme=3, Joint]"
           " [--forward] [--no-shot-gather] [--max-iter K] [--vmin V] [--vmax V]\n");
}

}  // namespace

int main(int argc, char **argv)
{
    std::vector<const char *> pos;
    bool forward = false, shot_gather = true;
    int MAXIter = 500;
    double vmin = 1600, vmax = 4500;
    for (int a = 1; a < argc; a++) {
        if (!strcmp(argv[a], "--forward")) forward = true;
        else if (!strcmp(argv[a], "--vmin") && a + 1 < argc) vmin = atof(argv[++a]);
        else if (!strcmp(argv[a], "--vmax") && a + 1 < argc) vmax = atof(argv[++a]);
        else if (!strcmp(argv[a], "--no-shot-gather")) shot_gather = false;
        else if (!strcmp(argv[a], "--max-iter") && a + 1 < argc) MAXIter = atoi(argv[++a]);
        else pos.push_back(argv[a]);
    }
    if (pos.empty() || pos.size() > 3) {
        usage();
        return 1;
    }
    const int LM = pos.size() >= 2 ? atoi(pos[1]) : 5;
    const int RTtag = pos.size() >= 3 ? atoi(pos[2]) : 3;
    printf(" * 2-D joint transmission / reflection traveltime tomography on the GPU (libpfast)\n");
    if (!forward) {
        printf("Optimization Method: L-BFGS\n%s Memory Number m=%d (Ignore if L-BFGS is not used)\n",
               pos.size() >= 2 ? "Specified" : "Default", LM);
        printf(" *****************************************************************\n");
        printf("%s\n", RTtag == 3 ? "Joint Inversion of Reflection and Transmission"
                       : RTtag == 1 ? "Inversion of Transmission ONLY"
                                    : "Inversion of Reflection ONLY");
        printf(" *****************************************************************\n");
        if (RTtag < 1 || RTtag > 3) {
            printf("RTtag is unrecognized!...abort...\n");
            return 1;
        }
    }

    // ---- parameter file (src/RT_PFAST2D.cpp:150-177)
    const std::vector<std::string> inp = drv::read_inp(pos[0]);
    const std::string vp0file = drv::tok(inp, 0, 0), srcfile = drv::tok(inp, 1, 0);
    const int sNX = atoi(drv::tok(inp, 2, 0).c_str()), sNY = atoi(drv::tok(inp, 2, 1).c_str());
    const std::string recfile = drv::tok(inp, 3, 0);
    const int rNX = atoi(drv::tok(inp, 4, 0).c_str()), rNY = atoi(drv::tok(inp, 4, 1).c_str());
    const std::string refgeom = drv::tok(inp, 5, 0);
    const int gmNX = atoi(drv::tok(inp, 6, 0).c_str()), gmNY = atoi(drv::tok(inp, 6, 1).c_str());
    Survey2D S;
    S.NX = atoi(drv::tok(inp, 7, 0).c_str());
    S.NY = atoi(drv::tok(inp, 7, 1).c_str());
    S.dx = atof(drv::tok(inp, 7, 2).c_str());
    S.dz = atof(drv::tok(inp, 7, 3).c_str());
    std::string vpfile, twfile;
    double nux = 80., nuy = 50.;
    int schemetag = 1, lsrc = 0;
    if (!forward) {
        vpfile = drv::tok(inp, 8, 0);
        nux = atof(drv::tok(inp, 9, 0).c_str());
        nuy = atof(drv::tok(inp, 9, 1).c_str());
        twfile = drv::tok(inp, 10, 0);
        schemetag = atoi(drv::tok(inp, 11, 0).c_str());
        lsrc = atoi(drv::tok(inp, 11, 1).c_str());
    }
    printf("\nThe initial velocity file... %s\n", vp0file.c_str());
    printf("The source location file... %s\n", srcfile.c_str());
    printf("The receiver location file... %s\n", recfile.c_str());
    printf("The source binary file dimension... %d x %d\n", sNX, sNY);
    printf("The receiver binary file dimension... %d x %d\n", rNX, rNY);
    printf("The reflector geometry file... %s\n", refgeom.c_str());
    printf("The reflector geometry binary file dimension... %d x %d\n", gmNX, gmNY);
    printf("The size of the model... %d X %d\nSample Interval... %.2lfm x %.2lfm\n", S.NX, S.NY,
           S.dx, S.dz);
    if (!forward) {
        printf("Final Velocity Model Output to %s\n", vpfile.c_str());
        printf("The weighting factor matrix file... %s\n\n", twfile.c_str());
    }

    // ---- inputs (src/RT_PFAST2D.cpp:225-404)
    S.NumofR = (int)ceil((double)gmNX / S.NY);
    const long N = (long)S.NX * S.NY;
    std::vector<double> srcArray, recArray;
    drv::readmod(S.vp, vp0file, N);
    drv::readmod(srcArray, srcfile, (size_t)sNX * sNY);
    drv::readmod(recArray, recfile, (size_t)rNX * rNY);
    drv::readmod(S.Z, refgeom, (size_t)gmNX * gmNY);
    S.NumofSrc = sNX;
    S.src.resize(2 * (size_t)S.NumofSrc);
    S.srnum.resize(S.NumofSrc);
    printf("...reading source file:\n\t%s\n ----------------------------------\n| Total Number of "
           "Sources: %d\n", srcfile.c_str(), S.NumofSrc);
    for (int ii = 0; ii < S.NumofSrc; ii++) {
        S.src[2 * ii + 0] = drv::map_index(srcArray[(size_t)ii * sNY + 0], S.dz);
        S.src[2 * ii + 1] = drv::map_index(srcArray[(size_t)ii * sNY + 1], S.dx);
        S.srnum[ii] = (int)(srcArray[(size_t)ii * sNY + 2]);
        if (ii <= 20 || ii == S.NumofSrc - 1)
            printf("| %d\t%d\t%d\t%d\n", ii, S.src[2 * ii] + 1, S.src[2 * ii + 1] + 1, S.srnum[ii]);
        else if (ii == 21)
            printf("| ......\n");
    }
    S.NumofRec = rNX;
    const int R = S.NumofR;
    S.rec.resize(2 * (size_t)rNX);
    S.nxy.resize(2 * (size_t)rNX);
    S.TT_obs.resize(rNX);
    S.TR_obs.resize((size_t)rNX * (R ? R : 1), 0.0);
    for (int ii = 0; ii < rNX; ii++) {
        const double *r = &recArray[(size_t)ii * rNY];
        S.rec[2 * ii] = drv::map_index(r[0], S.dz);
        S.rec[2 * ii + 1] = drv::map_index(r[1], S.dx);
        S.nxy[2 * ii] = rNY > 2 ? r[2] : 0.0;
        S.nxy[2 * ii + 1] = rNY > 3 ? r[3] : 0.0;
        // forward-modelling receiver files carry no picks (4 columns)
        S.TT_obs[ii] = rNY > 4 ? r[4] : 150.0;
        for (int it = 0; it < R; it++) S.TR_obs[(size_t)ii * R + it] = rNY > 5 + it ? r[5 + it] : 0.0;
    }
    long need = 0;
    for (int s = 0; s < S.NumofSrc; s++) need += S.srnum[s];
    if (need > rNX) {
        printf("the source table asks for %ld receiver rows, the receiver file has %d\n", need, rNX);
        return 1;
    }
    printf("...reading reflector geometry file:\n\t%s\n| Total Number of reflector geometry "
           "points: %d\n| Total Number of reflector layers: %d\n", refgeom.c_str(), gmNX, R);

    pfast_ctx *ctx = nullptr;
    check(pfast_ctx_create(0, &ctx), "pfast_ctx_create");
    pfast_model2d *model = nullptr;
    check(pfast_model2d_create(ctx, S.vp.data(), S.NX, S.NY, S.dx, S.dz, R ? S.Z.data() : nullptr, R,
                               &model),
          "pfast_model2d_create");
    const std::string p_vpfile = drv::prefix(vp0file, 3);

    if (forward) {
        // picks T and TTu per reflector for every receiver row, sources in order
        // (src/RT_PFAST2D.cpp:828-862), and the per-source fields
        pfast_shots2d *shots = nullptr;
        check(pfast_shots2d_create(model, S.NumofSrc, S.src.data(), S.srnum.data(), nullptr, nullptr,
                                   nullptr, nullptr, &shots),
              "pfast_shots2d_create");
        const int MAXSwp = 500;
        int rc = pfast_shots2d_run(shots, PFAST_FWD_ONLY, MAXSwp);
        if (rc == PFAST_ERR_FATAL) {
            printf("PE_0: %s...abort...\n", pfast_last_error());
            return 1;
        }
        check(rc, "pfast_shots2d_run");
        std::vector<double> Tr((size_t)need * (R + 1)), T(N), F((size_t)(R ? R : 1) * N);
        long chr = 0;
        for (int s = 0; s < S.NumofSrc; s++) {
            check(pfast_shots2d_field(shots, 0, s, 0, T.data()), "field");
            for (int idx = 0; idx < S.srnum[s]; idx++) {
                const int i = S.rec[2 * (idx + chr)], j = S.rec[2 * (idx + chr) + 1];
                Tr[(size_t)(chr + idx) * (R + 1)] = T[(long)i * S.NY + j];
            }
            for (int ir = 0; ir < R; ir++) {
                check(pfast_shots2d_field(shots, 2, s, ir, F.data() + (size_t)ir * N), "field");
                for (int idx = 0; idx < S.srnum[s]; idx++) {
                    const int i = S.rec[2 * (idx + chr)], j = S.rec[2 * (idx + chr) + 1];
                    Tr[(size_t)(chr + idx) * (R + 1) + ir + 1] = F[(size_t)ir * N + (long)i * S.NY + j];
                }
            }
            if (shot_gather && R) {
                drv::writemod(F.data(), p_vpfile + "_Src" + std::to_string(s) + "_TTu.2d", (size_t)R * N);
                for (int ir = 0; ir < R; ir++)
                    check(pfast_shots2d_field(shots, 1, s, ir, F.data() + (size_t)ir * N), "field");
                drv::writemod(F.data(), p_vpfile + "_Src" + std::to_string(s) + "_TTd.2d", (size_t)R * N);
            }
            chr += S.srnum[s];
        }
        drv::writemod(Tr.data(), p_vpfile + "_Tr2D.bin", Tr.size());
        pfast_shots2d_destroy(shots);
        pfast_model2d_destroy(model);
        pfast_ctx_destroy(ctx);
        return 0;
    }

    // ---- inversion (src/RT_PFAST2D.cpp:395-816)
    drv::readmod(S.WR, twfile, N);
    pfast_shots2d *shots = nullptr;
    check(pfast_shots2d_create(model, S.NumofSrc, S.src.data(), S.srnum.data(), S.rec.data(),
                               S.nxy.data(), S.TT_obs.data(), S.TR_obs.data(), &shots),
          "pfast_shots2d_create");
    const int MAXSwp = 500;
    drv::Optimiser opt;
    opt.init(ctx, N, LM);
    opt.schemetag = schemetag;
    opt.lsrc = lsrc;
    opt.MAXLineSch = 5;
    opt.dim2 = true;
    std::vector<double> global_E, global_Et, global_Er;
    std::vector<double> gT(N), gR(N), gRT(N);
    double E3[3];
    std::vector<double> vp = S.vp;
    const double NR = (double)S.NumofRec;

    // one batched gradient evaluation at `m`; Dj per RTtag, smoothed with the
    // current model `vp`; phi per RTtag
    auto evaluate = [&](const std::vector<double> &m) {
        check(pfast_model2d_set_vp(model, m.data()), "set_vp");
        int rc = pfast_shots2d_run(shots, PFAST_FWD_ADJ, MAXSwp);
        if (rc == PFAST_ERR_FATAL) {
            printf("PE_0: %s...abort...\n", pfast_last_error());
            exit(1);
        }
        check(rc, "pfast_shots2d_run");
        check(pfast_shots2d_gradient(shots, S.WR.data(), gT.data(), gR.data(), gRT.data(), E3),
              "pfast_shots2d_gradient");
    };
    drv::EvalFn trial = [&](const std::vector<double> &m, std::vector<double> &Dj, double &phi) {
        evaluate(m);
        Dj = RTtag == 3 ? gRT : RTtag == 1 ? gT : gR;
        phi = RTtag == 3 ? E3[2] : RTtag == 1 ? E3[0] : E3[1];
        check(pfast_regu2d(Dj.data(), vp.data(), nux, nuy, S.dx, S.dz, S.NX, S.NY), "regu2d");
    };

    bool convflag = false;
    int Iterk = 0;
    while (!convflag && Iterk++ < MAXIter) {
        printf("Iterating No. %d: (nux,nuy)=(%f,%f)\n", Iterk, nux, nuy);
        opt.Djk0 = opt.Djk1;
        evaluate(vp);
        opt.Djk1 = RTtag == 3 ? gRT : RTtag == 1 ? gT : gR;
        global_E.push_back(E3[2]);
        global_Et.push_back(E3[0]);
        global_Er.push_back(E3[1]);
        printf("Average Time Residual (TR): %g ms\t(T): %g ms\t(R): %g ms\n",
               1000 * sqrt(2 * E3[2] / NR), 1000 * sqrt(2 * E3[0] / NR), 1000 * sqrt(2 * E3[1] / NR));
        const std::string k = std::to_string(Iterk);
        drv::writemod(opt.Djk1.data(), p_vpfile + "_bf_regu_RT" + k, N);
        check(pfast_regu2d(opt.Djk1.data(), vp.data(), nux, nuy, S.dx, S.dz, S.NX, S.NY), "regu2d");
        if (schemetag != 3 || Iterk == 1) opt.alpha1 = 5e2 * pfast_magnitude_of(opt.Djk1.data(), N);
        drv::writemod(opt.Djk1.data(), p_vpfile + "_af_regu_RT" + k, N);
        drv::writemod(gR.data(), p_vpfile + "_bf_regu_R" + k, N);
        check(pfast_regu2d(gR.data(), vp.data(), nux, nuy, S.dx, S.dz, S.NX, S.NY), "regu2d");
        drv::writemod(gR.data(), p_vpfile + "_af_regu_R" + k, N);
        drv::writemod(gT.data(), p_vpfile + "_bf_regu_T" + k, N);
        check(pfast_regu2d(gT.data(), vp.data(), nux, nuy, S.dx, S.dz, S.NX, S.NY), "regu2d");
        drv::writemod(gT.data(), p_vpfile + "_af_regu_T" + k, N);
        {
            FILE *fe = fopen((p_vpfile + "_GlobalErr_TR_T_R.txt").c_str(), "w");
            if (fe) {
                for (int i = 0; i < Iterk; i++)
                    fprintf(fe, "%g\t%g\t%g\n", global_E[i], global_Et[i], global_Er[i]);
                fclose(fe);
            }
        }
        if (!opt.direction(Iterk)) {
            printf("Numerical Optimization Method Unknown..abort...\n");
            return 1;
        }
        opt.line_search(vp, global_E[Iterk - 1], trial);
        drv::updatevp(vp, opt.dk, opt.alpha1, vmin, vmax);
        drv::writemod(vp.data(), p_vpfile + "_" + k, N);
        convflag = drv::check_conv(opt.dk, opt.alpha1, vp, global_E, Iterk, 0.001, N);
        printf("--------Interating No. %d done---------------\n\n", Iterk);
    }
    drv::writemod(vp.data(), vpfile, N);
    {
        FILE *fe = fopen((p_vpfile + "_AveRes_TR_T_R.txt").c_str(), "w");
        if (fe) {
            for (size_t i = 0; i < global_E.size(); i++)
                fprintf(fe, "%g\t%g\t%g\n", 1000 * sqrt(2 * global_E[i] / NR),
                        1000 * sqrt(2 * global_Et[i] / NR), 1000 * sqrt(2 * global_Er[i] / NR));
            fclose(fe);
        }
    }
    pfast_shots2d_destroy(shots);
    pfast_model2d_destroy(model);
    pfast_ctx_destroy(ctx);
    return 0;
}
