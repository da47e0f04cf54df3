// slab_schedule.cpp -- CPU model of the batched 3D "slab ring" sweep schedule
// (pfast_amd/csrc/batch3d.hip: fwd3d_slab / adj3d_slab), one source.
//
// Test infrastructure: it replays the kernels' schedule exactly -- tiles launched
// by anti-diagonal, per-thread levels pk = p - (pi+pj), the W-slot LDS ring with
// its incremental slot counters and sentinel levels, the L-deep register
// prefetch, the retire-and-write-back of changed nodes -- with every thread of a
// block run between two barriers in turn, and checks that one ordering gives the
// same bits as the plain lexicographic sweep of the same ordering (the
// reference's loop order) with the same per-node arithmetic.
//
// Built with -DPLANT_NO_SKEW or -DPLANT_RING_DEPTH it carries a deliberate
// schedule mistake, which the tests expect it to report.
//
// usage: slab_schedule <fwd|adj> nrows ncols nz TI TJ L input.bin
//   input.bin: float T[N] (adj: traveltimes, negative = fixed node; fwd: cell
//   time h/v), float U0[N] (initial field), int src (fwd: source node or -1)
// prints "orderings <n> mismatches <m>" and exits 1 on any mismatch.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

static const float INFF = 9999999.0f;
static int nr, nc, nz;

static float eik_uni(float a, float b, float c, float f)
{
    const float lo = fminf(fminf(a, b), c), hi = fmaxf(fmaxf(a, b), c);
    const float mid = fmaxf(fminf(a, b), fminf(fmaxf(a, b), c));
    const float q1 = lo + f, d = mid - lo;
    const float q2 = mid + 0.5f * (sqrtf(fmaf(-d, d, 2.0f * f * f)) - d);
    const float db = hi - mid, dc = hi - lo, P = db + dc;
    const float disc3 = fmaf(P, P, -3.0f * (fmaf(db, db, dc * dc) - f * f));
    const float q3 = disc3 >= 0.0f ? hi + (sqrtf(disc3) - P) * (1.0f / 3.0f) : INFF;
    return q1 <= mid ? q1 : (q2 <= hi ? q2 : q3);
}

static float term(float tc, float tn, float ln, float w, float &den)
{
    den = fmaf(w, fmaxf(tc - tn, 0.0f), den);
    return tn > tc ? w * (tn - tc) * ln : 0.0f;
}

static float adj(float tc, float txm, float txp, float tym, float typ, float tzm, float tzp,
                 float lxm, float lxp, float lym, float lyp, float lzm, float lzp)
{
    const float w = 1.0f;
    if (tc >= INFF || txm >= INFF || txp >= INFF || tym >= INFF || typ >= INFF || tzm >= INFF ||
        tzp >= INFF)
        return 0.0f;
    float den = 0.0f;
    float num = term(tc, txp, lxp, w, den);
    num += term(tc, txm, lxm, w, den);
    num += term(tc, typ, lyp, w, den);
    num += term(tc, tym, lym, w, den);
    num += term(tc, tzp, lzp, w, den);
    num += term(tc, tzm, lzm, w, den);
    return fabsf(den) > 1e-9f ? num / den : 0.0f;
}

static void lexi(bool fwd, const float *T, float *U, int src, int o)
{
    const long plane = (long)nr * nc;
    const int s3 = o < 4 ? -1 : 1, s1 = (o & 2) ? 1 : -1, s2 = (o & 1) ? 1 : -1;
    const int b = fwd ? 0 : 1;
    for (int kk = b; kk < nz - b; kk++) {
        const int k = s3 < 0 ? nz - 1 - kk : kk;
        for (int ii = b; ii < nr - b; ii++) {
            const int i = s1 < 0 ? nr - 1 - ii : ii;
            for (int jj = b; jj < nc - b; jj++) {
                const int j = s2 < 0 ? nc - 1 - jj : jj;
                const long n = k * plane + (long)i * nc + j;
                if (fwd) {
                    if (n == src) continue;
                    auto at = [&](long m, bool in) { return in ? U[m] : INFF; };
                    const float xa = fminf(at(n - 1, j > 0), at(n + 1, j < nc - 1));
                    const float ya = fminf(at(n - nc, i > 0), at(n + nc, i < nr - 1));
                    const float za = fminf(at(n - plane, k > 0), at(n + plane, k < nz - 1));
                    const float q = eik_uni(xa, ya, za, T[n]);
                    if (q < U[n]) U[n] = q;
                } else {
                    if (std::signbit(T[n])) continue;
                    U[n] = adj(T[n], fabsf(T[n - 1]), fabsf(T[n + 1]), fabsf(T[n - nc]),
                               fabsf(T[n + nc]), fabsf(T[n - plane]), fabsf(T[n + plane]), U[n - 1],
                               U[n + 1], U[n - nc], U[n + nc], U[n - plane], U[n + plane]);
                }
            }
        }
    }
}

// One ordering with the kernels' schedule; G is the field in "global memory".
static void slab(bool fwd, const float *T, float *G, int src, int o, int TI, int TJ, int L)
{
    const int s3 = o < 4 ? -1 : 1, s1 = (o & 2) ? 1 : -1, s2 = (o & 1) ? 1 : -1;
    const int HJ = TJ + 2, SLAB = (TI + 2) * HJ, W = TI + TJ + 3;
    const unsigned plane = (unsigned)nr * nc;
    const int nti = (nr + TI - 1) / TI, ntj = (nc + TJ - 1) / TJ;
    const int NT = TI * TJ, NH = 2 * (TI + TJ);
    for (int D = 0; D <= nti + ntj - 2; D++) {
        const int lo = D - (ntj - 1) > 0 ? D - (ntj - 1) : 0, hi = D < nti - 1 ? D : nti - 1;
        for (int ti_ = lo; ti_ <= hi; ti_++) {
            const int tj_ = D - ti_;
            const int ti = s1 > 0 ? ti_ : nti - 1 - ti_, tj = s2 > 0 ? tj_ : ntj - 1 - tj_;
            const int i0 = ti * TI, j0 = tj * TJ;
            const int TIe = TI < nr - i0 ? TI : nr - i0, TJe = TJ < nc - j0 ? TJ : nc - j0;
            const int span = TIe + TJe - 2;
            // ring of field values (fwd: plus cell times) and, for adj, traveltimes
            std::vector<float> ring(W * SLAB, NAN), tr(W * SLAB, NAN), fr(W * NT, NAN);
            struct Th {
                bool valid, huse, hreal;
                int li, lj, pipj, me, hl;
                unsigned oI, oH, oS, oF;
                int srclev;
                unsigned dirty;
                std::vector<float> rI, rH, rT, rTH, rF;
            };
            std::vector<Th> th(NT);
            const unsigned k0 = s3 > 0 ? 0u : (unsigned)(nz - 1);
            const int step = s3 > 0 ? (int)plane : -(int)plane;
            const unsigned org = (unsigned)i0 * nc + j0;
            for (int t = 0; t < NT; t++) {
                Th &h = th[t];
                h.li = t / TJ;
                h.lj = t % TJ;
                h.valid = h.li < TIe && h.lj < TJe;
                h.pipj = (s1 > 0 ? h.li : TIe - 1 - h.li) + (s2 > 0 ? h.lj : TJe - 1 - h.lj);
#ifdef PLANT_NO_SKEW  // planted mistake (tests only): every column on the same level
                h.pipj = 0;
#endif
                h.me = (h.li + 1) * HJ + h.lj + 1;
                const unsigned node = h.valid ? (unsigned)(i0 + h.li) * nc + j0 + h.lj : org;
                // halo element t (side list order: top, bottom, left, right)
                h.huse = false;
                h.hreal = false;
                h.hl = 0;
                unsigned hnode = org;
                if (t < NH) {
                    int q = t, hi_ = 0, hj_ = 0, gi = 0, gj = 0;
                    bool ok = true;
                    if (q < TJ) { ok = q < TJe; hi_ = 0; hj_ = q + 1; gi = i0 - 1; gj = j0 + q; }
                    else if (q < 2 * TJ) { q -= TJ; ok = q < TJe; hi_ = TIe + 1; hj_ = q + 1; gi = i0 + TIe; gj = j0 + q; }
                    else if (q < 2 * TJ + TI) { q -= 2 * TJ; ok = q < TIe; hi_ = q + 1; hj_ = 0; gi = i0 + q; gj = j0 - 1; }
                    else { q -= 2 * TJ + TI; ok = q < TIe; hi_ = q + 1; hj_ = TJe + 1; gi = i0 + q; gj = j0 + TJe; }
                    if (ok) {
                        h.huse = true;
                        h.hreal = gi >= 0 && gi < nr && gj >= 0 && gj < nc;
                        h.hl = hi_ * HJ + hj_;
                        if (h.hreal) hnode = (unsigned)gi * nc + gj;
                    }
                }
                h.oI = k0 * plane + node;
                h.oH = k0 * plane + hnode;
                h.oS = h.oI;
                h.oF = h.oI;
                h.srclev = -1;
                if (fwd && h.valid && src >= 0 && (unsigned)src % plane == node) {
                    const int sk = (int)((unsigned)src / plane);
                    h.srclev = s3 > 0 ? sk : nz - 1 - sk;
                }
                h.dirty = 0;
                h.rI.assign(L, 0);
                h.rH.assign(L, 0);
                h.rT.assign(L, 0);
                h.rTH.assign(L, 0);
                h.rF.assign(L, 0);
            }
            auto put = [&](Th &h, int t, int slot, float vi, float vh, float vt, float vth, float vf) {
                if (h.valid) { ring[slot * SLAB + h.me] = vi; tr[slot * SLAB + h.me] = vt; fr[slot * NT + t] = vf; }
                if (h.huse) {
                    ring[slot * SLAB + h.hl] = fwd ? (h.hreal ? vh : INFF) : vh;
                    tr[slot * SLAB + h.hl] = h.hreal ? vth : INFF;
                }
            };
            for (int t = 0; t < NT; t++) {
                Th &h = th[t];
                if (fwd && h.valid) ring[(W - 1) * SLAB + h.me] = INFF;
                put(h, t, 0, G[h.oI], G[h.oH], T[h.oI], T[h.oH], T[h.oF]);
                for (int u = 0; u < L; u++)
                    if (u + 1 < nz) {
                        h.oI += step; h.oH += step; h.oF += step;
                        h.rI[u] = G[h.oI]; h.rH[u] = G[h.oH]; h.rT[u] = T[h.oI]; h.rTH[u] = T[h.oH]; h.rF[u] = T[h.oF];
                    }
            }
            const int pend = nz + span + 2;
            std::vector<int> vs(NT);
            for (int t = 0; t < NT; t++) vs[t] = th[t].pipj == 0 ? 0 : W - th[t].pipj;
            int ss = 1, sr = W - (span + 2);
            for (int pb = 0; pb < pend; pb += L)
                for (int u = 0; u < L; u++) {
                    const int p = pb + u;
                    for (int t = 0; t < NT; t++) {  // steps (a), (b)
                        Th &h = th[t];
                        if (p + 1 < nz) put(h, t, ss, h.rI[u], h.rH[u], h.rT[u], h.rTH[u], h.rF[u]);
                        else if (fwd && p + 1 == nz && h.valid) ring[ss * SLAB + h.me] = INFF;
                        if (p + 1 + L < nz) {
                            h.oI += step; h.oH += step; h.oF += step;
                            h.rI[u] = G[h.oI]; h.rH[u] = G[h.oH]; h.rT[u] = T[h.oI]; h.rTH[u] = T[h.oH]; h.rF[u] = T[h.oF];
                        }
                    }
                    // barrier
                    for (int t = 0; t < NT; t++) {  // steps (d), (e)
                        Th &h = th[t];
                        const int pk = p - h.pipj;
                        const int v = vs[t], vm = v == 0 ? W - 1 : v - 1, vp = v == W - 1 ? 0 : v + 1;
                        const int o0 = v * SLAB + h.me;
                        if (fwd) {
                            if (h.valid && pk >= 0 && pk < nz) {
                                const float self = ring[o0];
                                const float xa = fminf(ring[o0 - 1], ring[o0 + 1]);
                                const float ya = fminf(ring[o0 - HJ], ring[o0 + HJ]);
                                const float za = fminf(ring[vm * SLAB + h.me], ring[vp * SLAB + h.me]);
                                const float q = eik_uni(xa, ya, za, fr[v * NT + h.li * TJ + h.lj]);
                                if (q < self && pk != h.srclev) { ring[o0] = q; h.dirty |= 1u << (pk & 31); }
                            }
                        } else {
                            const int gi = i0 + h.li, gj = j0 + h.lj;
                            const bool ci = h.valid && gi >= 1 && gi <= nr - 2 && gj >= 1 && gj <= nc - 2;
                            if (ci && pk >= 1 && pk <= nz - 2 && !std::signbit(tr[o0])) {
                                const int lo_ = vm * SLAB + h.me, hi2 = vp * SLAB + h.me;
#ifndef PLANT_RING_DEPTH  // planted mistake (tests only): name z-neighbours by ring order
                                const int om = s3 > 0 ? lo_ : hi2, op = s3 > 0 ? hi2 : lo_;
#else
                                const int om = lo_, op = hi2;
#endif
                                const float nw = adj(tr[o0], fabsf(tr[o0 - 1]), fabsf(tr[o0 + 1]),
                                                     fabsf(tr[o0 - HJ]), fabsf(tr[o0 + HJ]), fabsf(tr[om]),
                                                     fabsf(tr[op]), ring[o0 - 1], ring[o0 + 1], ring[o0 - HJ],
                                                     ring[o0 + HJ], ring[om], ring[op]);
                                if (nw != ring[o0]) { ring[o0] = nw; h.dirty |= 1u << (pk & 31); }
                            }
                        }
                        vs[t] = v == W - 1 ? 0 : v + 1;
                        const int lr = p - span - 2;
                        if (lr >= 0 && lr < nz) {
                            if (h.valid && ((h.dirty >> (lr & 31)) & 1u)) {
                                G[h.oS] = ring[sr * SLAB + h.me];
                                h.dirty &= ~(1u << (lr & 31));
                            }
                            h.oS += step;
                        }
                    }
                    sr = sr == W - 1 ? 0 : sr + 1;
                    ss = ss == W - 1 ? 0 : ss + 1;
                }
        }
    }
}

int main(int argc, char **argv)
{
    if (argc < 9) return 2;
    const bool fwd = argv[1][0] == 'f';
    nr = atoi(argv[2]);
    nc = atoi(argv[3]);
    nz = atoi(argv[4]);
    const int TI = atoi(argv[5]), TJ = atoi(argv[6]), L = atoi(argv[7]);
    const long N = (long)nr * nc * nz;
    std::vector<float> T(N), U(N);
    int src = -1;
    FILE *f = fopen(argv[8], "rb");
    if (!f || fread(T.data(), 4, N, f) != (size_t)N || fread(U.data(), 4, N, f) != (size_t)N ||
        fread(&src, 4, 1, f) != 1)
        return 3;
    fclose(f);
    std::vector<float> A = U, B = U;
    long bad = 0;
    int n = 0;
    for (int cyc = 0; cyc < 2; cyc++)
        for (int o = 0; o < 8; o++, n++) {
            lexi(fwd, T.data(), A.data(), src, o);
            slab(fwd, T.data(), B.data(), src, o, TI, TJ, L);
            for (long m = 0; m < N; m++)
                if (A[m] != B[m] && !(std::isnan(A[m]) && std::isnan(B[m]))) bad++;
            B = A;  // continue both from the same state
        }
    printf("orderings %d mismatches %ld\n", n, bad);
    return bad ? 1 : 0;
}
