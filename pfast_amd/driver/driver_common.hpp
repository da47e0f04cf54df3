// driver_common.hpp -- pieces shared by the 2D and 3D tomography drivers
// (pfast_rt2d, pfast_t3d): parameter files, raw model files, and the optimiser
// (search directions and line searches) of the reference drivers
// src/RT_PFAST2D.cpp:412-816 and src/T_PFAST3D.cpp:359-655, whose gradient
// evaluations here are one batched GPU call instead of a per-source loop.
//
// The L-BFGS memory, its two-loop recursion and the line searches' dot products run
// on the GPU in FP64 with compensated sums (pfast_lbfgs_*); the conjugate-gradient
// betas and the clamped update stay host loops in the reference's order.
#pragma once
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "../../include/pfast.h"

namespace drv {

// ------------------------------------------------------------------ files
// Parameter file: every line that does not start with '#' is the next parameter
// line, as src/RT_PFAST2D.cpp:140-148 reads it.
inline std::vector<std::string> read_inp(const char *path)
{
    std::vector<std::string> out;
    FILE *fp = fopen(path, "r");
    if (!fp) {
        printf("The file %s was not opened\n", path);
        exit(1);
    }
    printf("Read input parameters from file %s\nUser Parameters are as follows:\n", path);
    char line[1024];
    while (fgets(line, sizeof line, fp))
        if (line[0] != '#') out.emplace_back(line);
    fclose(fp);
    return out;
}

// Whitespace-separated tokens of a parameter line.
inline std::vector<std::string> tokens(const std::string &line)
{
    std::vector<std::string> t;
    size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && isspace((unsigned char)line[i])) i++;
        size_t j = i;
        while (j < line.size() && !isspace((unsigned char)line[j])) j++;
        if (j > i) t.push_back(line.substr(i, j - i));
        i = j;
    }
    return t;
}

inline std::string tok(const std::vector<std::string> &inp, size_t line, size_t k)
{
    if (line >= inp.size()) {
        printf("parameter file: line %zu missing\n", line + 1);
        exit(1);
    }
    std::vector<std::string> t = tokens(inp[line]);
    if (k >= t.size()) {
        printf("parameter file: field %zu of parameter line %zu missing\n", k + 1, line + 1);
        exit(1);
    }
    return t[k];
}

// Raw little-endian FP64 arrays (the reference's readmod / writemod).
inline void readmod(std::vector<double> &a, const std::string &file, size_t n)
{
    FILE *fp = fopen(file.c_str(), "rb");
    if (!fp) {
        printf(" The file %s cannot be opened\n", file.c_str());
        exit(1);
    }
    printf(" \n...reading file:\n\t %s\n", file.c_str());
    a.assign(n, 0.0);
    if (fread(a.data(), sizeof(double), n, fp) != n) printf(" short read from %s\n", file.c_str());
    fclose(fp);
}

inline void writemod(const double *a, const std::string &file, size_t n)
{
    FILE *fp = fopen(file.c_str(), "wb");
    if (!fp) {
        printf(" The file %s cannot be opened\n", file.c_str());
        return;
    }
    printf(" \n...writing file:\n\t %s\n", file.c_str());
    fwrite(a, sizeof(double), n, fp);
    fclose(fp);
}

// Coordinate -> node index, src/T_PFAST3D.cpp:250-252: node i sits at (i+1) h.
inline int map_index(double c, double h) { return (int)ceil(c / h) - 1; }

inline void check(int rc, const char *what)
{
    if (rc == PFAST_OK) return;
    printf("%s failed: %s\n", what, pfast_last_error());
    exit(1);
}

// ------------------------------------------------------------------ optimiser
// One gradient evaluation at a trial model: the regularised gradient (Dj, N
// doubles) and the misfit phi.
using EvalFn = std::function<void(const std::vector<double> &vp_trial, std::vector<double> &Dj,
                                  double &phi)>;

struct Optimiser {
    long N = 0;
    int LM = 5;
    int schemetag = 3, lsrc = 2, MAXLineSch = 5;
    bool dim2 = true;  // 2D or 3D flavour of the small differences below
    double alpha0 = 0.0, alpha1 = 5.0e3, c1 = 1.0e-4, c2 = 0.9;
    std::vector<double> Djk1, Djk0, dk, Dj1, loc_vp;
    pfast_lbfgs *lb = nullptr;  // the L-BFGS pairs on the GPU (sk, yk, rhoi of the reference)

    void init(pfast_ctx *ctx, long n, int lm)
    {
        N = n;
        LM = lm;
        Djk1.assign(N, 0.0);
        Djk0.assign(N, 0.0);
        dk.assign(N, 0.0);
        Dj1.assign(N, 0.0);
        loc_vp.assign(N, 0.0);
        check(pfast_lbfgs_create(ctx, N, LM, &lb), "pfast_lbfgs_create");
    }

    ~Optimiser() { pfast_lbfgs_destroy(lb); }

    // FP64 compensated a . b on the GPU
    double dot(const std::vector<double> &a, const std::vector<double> &b) const
    {
        double d = 0.0;
        check(pfast_lbfgs_dot(lb, a.data(), b.data(), &d), "pfast_lbfgs_dot");
        return d;
    }

    // Search direction at iteration Iterk (src/RT_PFAST2D.cpp:500-552).  Returns
    // false for an unknown scheme (the reference aborts).  The L-BFGS recursion
    // reads pair i from slot i % LM: the reference reads ai[i], which leaves the
    // array once Iterk >= LM + 2 (src/RT_PFAST2D.cpp:529, src/T_PFAST3D.cpp:467).
    bool direction(int Iterk)
    {
        if (Iterk == 1) {
            for (long i = 0; i < N; i++) dk[i] = Djk1[i];
            return true;
        }
        switch (schemetag) {
        case 1:
        case 2: {
            const double beta = schemetag == 1 ? pfast_beta_hs(Djk1.data(), Djk0.data(), dk.data(), N)
                                               : pfast_beta_cgdescent(Djk1.data(), Djk0.data(),
                                                                      dk.data(), N);
            for (long i = 0; i < N; i++) dk[i] = Djk1[i] + dk[i] * beta;
            c1 = 1.0e-4;
            c2 = 0.1;
            return true;
        }
        case 3: {
            check(pfast_lbfgs_direction(lb, Iterk, Djk1.data(), Djk0.data(), dk.data(), alpha1,
                                        nullptr),
                  "pfast_lbfgs_direction");
            alpha0 = 0.0;
            alpha1 = 1.0;
            c1 = 1.0e-4;
            c2 = 0.9;
            return true;
        }
        default:
            return false;
        }
    }

    // dphi = -dk . g (src/RT_PFAST2D.cpp:622)
    double dphi(const std::vector<double> &g) const { return -dot(dk, g); }

    // Line search from vp along dk (src/RT_PFAST2D.cpp:561-792); phi0 is the misfit
    // of the current model.  On return alpha1 is the step.  `eval` regularises
    // with the current model.  `on_alpha` (3D, strong-Wolfe secant) receives the
    // final step (the reference writes it to alpha1.tmp).
    void line_search(const std::vector<double> &vp, double phi0, const EvalFn &eval,
                     const std::function<void(double)> &on_alpha = nullptr)
    {
        int count = 0;
        switch (lsrc) {
        case 1: {  // secant on dk . g = 0 ("exact" condition)
            printf("Secant Method Line Search for 'exact' condition (alpha1=%g)...\n", alpha1);
            alpha0 = 0.0;
            double f0 = 0.0, f1;
            while (fabs(alpha1 - alpha0) / fabs(alpha0) > 1e-2 && count++ < MAXLineSch) {
                for (long i = 0; i < N; i++) loc_vp[i] = vp[i] + alpha1 * dk[i];
                double phi1;
                eval(loc_vp, Dj1, phi1);
                // the reference zeroes f0 right before the update (its running f0 of
                // the previous trial is never used)
                f0 = 0.0;
                f1 = dot(Dj1, dk);
                const double alphak = alpha1 - (alpha1 - alpha0) * f1 / (f1 - f0);
                alpha0 = alpha1;
                alpha1 = alphak;
                f0 = f1;
            }
            if (count > MAXLineSch)
                printf("Line Search Stopped, Stepsize=%f\n ...Updating velocity now...\n", alpha1);
            else
                printf("Line Search Completed After %d Iterations, Stepsize=%f\n...Updating "
                       "velocity now...\n", count, alpha1);
            break;
        }
        case 2:
        case 3: {
            const bool cubic = lsrc == 3;
            if (cubic)
                printf("Cubic Interpolation Line Search with Strong Wolfe Condition (alpha1=%g)...\n",
                       alpha1);
            else
                printf("Secant Method Line Search with Strong Wolfe Condition (alpha1=%g)...\n",
                       alpha1);
            bool isWolfe = false;
            double phi1 = 0.0, dphi1 = 0.0;
            alpha0 = 0.0;
            const double dphi0 = dphi(Djk1);
            // 2D with a conjugate-gradient scheme: one cubic step, no Wolfe test
            // (src/RT_PFAST2D.cpp:757-790)
            const bool single = cubic && dim2 && schemetag != 3;
            while (!isWolfe && count++ < MAXLineSch) {
                for (long i = 0; i < N; i++) loc_vp[i] = vp[i] + alpha1 * dk[i];
                eval(loc_vp, Dj1, phi1);
                dphi1 = dphi(Dj1);
                if (single) {
                    alpha1 = pfast_cubic_interp(alpha0, alpha1, phi0, phi1, dphi0, dphi1);
                    printf("(alpha0,alpha1,phi0,phi1,dphi0,dphi1)=(%g %g %g %g %g %g\n", alpha0,
                           alpha1, phi0, phi1, dphi0, dphi1);
                    break;
                }
                if (!cubic)
                    printf("(alpha0,alpha1,phi0,phi1,dphi0,dphi1)=(%g %g %g %g %g %g\n", alpha0,
                           alpha1, phi0, phi1, dphi0, dphi1);
                if (phi1 <= phi0 + c1 * alpha1 * dphi0 && fabs(dphi1) <= -c2 * dphi0)
                    isWolfe = true;
                else if (cubic)
                    alpha1 = pfast_cubic_interp(alpha0, alpha1, phi0, phi1, dphi0, dphi1);
                else
                    alpha1 = alpha1 - (alpha1 - alpha0) * dphi1 / (dphi1 - dphi0);
                if (cubic)
                    printf("(alpha0,alpha1,phi0,phi1,dphi0,dphi1)=(%g %g %g %g %g %g\n", alpha0,
                           alpha1, phi0, phi1, dphi0, dphi1);
            }
            if (!cubic && on_alpha) on_alpha(alpha1);
            if (!single) {
                if (count > MAXLineSch)
                    printf("Line Search Stopped, Stepsize=%f\n ...Updating velocity now...\n",
                           alpha1);
                else
                    printf("Wolfe condition satisfied After %d Iterations, Stepsize=%f\n...Updating "
                           "velocity now...\n", count, alpha1);
            }
            break;
        }
        case 0:
            printf("No Line Searching, Stepsize=%f\n", alpha1);
            break;
        default:
            printf("Line search tag unknown..abort...\n");
            exit(1);
        }
    }
};

// Clamped update of non-air nodes (src/RT_padjoint2d.h:496-504, src/T_padjoint3d.h:195-203).
inline void updatevp(std::vector<double> &vp, const std::vector<double> &dm, double stepsize,
                     double vmin, double vmax)
{
    for (size_t k = 0; k < vp.size(); k++) {
        if (fabs(vp[k] - 330) > 1e-9) {
            vp[k] = vp[k] + stepsize * dm[k];
            if (vp[k] > vmax) vp[k] = vmax;
            if (vp[k] < vmin) vp[k] = vmin;
        }
    }
}

// Relative misfit decrease test (src/RT_padjoint2d.h:475-494: threshold 1e-3 in
// 2D; src/T_padjoint3d.h:174-193: 1e-2 in 3D) and the over-correction abort.
inline bool check_conv(const std::vector<double> &dvp, double stepsize, const std::vector<double> &vp,
                       const std::vector<double> &dEE, int Iterk, double threshold, long ncells)
{
    const double chgE = Iterk >= 2 ? (dEE[Iterk - 2] - dEE[Iterk - 1]) / dEE[Iterk - 2] : 1;
    double mynorm = 0;
    for (size_t k = 0; k < vp.size(); k++) {
        const double v = stepsize * stepsize * dvp[k] * dvp[k] / vp[k] / vp[k];
        if (v >= mynorm) mynorm = v;
    }
    mynorm = sqrt(mynorm);
    if (chgE <= threshold) return true;
    if (mynorm / (double)ncells >= 1) {
        printf("Over corrected..abort...\n");
        exit(1);
    }
    return false;
}

// Reference-order FP64 source/receiver mapping helpers
inline std::string prefix(const std::string &vp0file, size_t strip)
{
    return vp0file.size() > strip ? vp0file.substr(0, vp0file.size() - strip) : std::string();
}

}  // namespace drv
