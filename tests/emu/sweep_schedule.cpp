// sweep_schedule.cpp -- CPU model of the batched 3D sweep schedule of
// pfast_amd/csrc/batch3d.hip (the column window, fwd3d_col / adj3d_col), one source.
//
// Test infrastructure: it replays the kernels' schedule exactly -- tiles launched
// by anti-diagonal, per-thread levels pk = p - (pi+pj), the per-step publication
// of levels m-1 / m+1 into the parity-buffered exchange array, the neighbour reads
// (upstream "m-1", downstream "m+1", halo or INF slot at tile edges), the batches
// of L steps whose loads are issued one batch ahead (reads outside the column
// return 0, as the kernels' buffer loads do) and whose changed nodes are stored at
// the start of the next batch -- with every thread of a block run between two
// barriers in turn, and checks that one ordering gives the same bits as the plain
// lexicographic sweep of the same ordering (the reference's loop order) with the
// same per-node arithmetic.
//
// Built with -DPLANT_NO_SKEW, -DPLANT_RING_DEPTH or -DPLANT_SWAP_EXCHANGE it
// carries a deliberate schedule mistake, which the tests expect it to report.
//
// usage: sweep_schedule <fwd|adj> nrows ncols nz TI TJ L input.bin
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

// One ordering with the column-window schedule; G is the field in HBM.
static void col(bool fwd, const float *T, float *G, int src, int o, int TI, int TJ, int L)
{
    const int s3 = o < 4 ? -1 : 1, s1 = (o & 2) ? 1 : -1, s2 = (o & 1) ? 1 : -1;
    const long plane = (long)nr * nc;
    const int nti = (nr + TI - 1) / TI, ntj = (nc + TJ - 1) / TJ;
    const int NT = TI * TJ, INFSLOT = 2 * NT;
    // value of level lev (sweep order) of column (i,j); 0 outside the column
    auto ld = [&](const float *F, int i, int j, int lev) -> float {
        if (lev < 0 || lev >= nz) return 0.0f;
        const int k = s3 > 0 ? lev : nz - 1 - lev;
        return F[k * plane + (long)i * nc + j];
    };
    auto stv = [&](int i, int j, int lev, float v) {
        const int k = s3 > 0 ? lev : nz - 1 - lev;
        G[k * plane + (long)i * nc + j] = v;
    };
    for (int D = 0; D <= nti + ntj - 2; D++) {
        const int lo = D - (ntj - 1) > 0 ? D - (ntj - 1) : 0, hi = D < nti - 1 ? D : nti - 1;
        for (int ti_ = lo; ti_ <= hi; ti_++) {
            const int tj_ = D - ti_;
            const int ti = s1 > 0 ? ti_ : nti - 1 - ti_, tj = s2 > 0 ? tj_ : ntj - 1 - tj_;
            const int i0 = ti * TI, j0 = tj * TJ;
            const int TIe = TI < nr - i0 ? TI : nr - i0, TJe = TJ < nc - j0 ? TJ : nc - j0;
            const int span = TIe + TJe - 2;
            struct V { float l, t; };  // field value, traveltime (adjoint)
            struct Th {
                bool valid, uxm, uxp, uym, uyp, inter;
                int li, lj, pipj, rxm, rxp, rym, ryp, gi, gj, hxi, hxj, hyi, hyj, plo, phi, psrc;
                V prev, cur, next;
                std::vector<V> q[3], nq[3];  // own (m+2), x halo, y halo
                std::vector<float> fq, nfq, sval;
                std::vector<int> chg;
            };
            std::vector<Th> th(NT);
            const int PB = 2 * NT + 1;  // per step parity: [m-1 | m+1][thread], INF slot
            std::vector<V> pub(2 * PB);
            pub[INFSLOT] = pub[PB + INFSLOT] = V{fwd ? INFF : 0.0f, INFF};
            auto loadv = [&](int i, int j, int lev) {
                return V{ld(G, i, j, lev), fwd ? 0.0f : ld(T, i, j, lev)};
            };
            auto load_batch = [&](Th &h, int pb, std::vector<V> *q, std::vector<float> &fv) {
                const int mb = pb - h.pipj;
                for (int u = 0; u < L; u++) {
                    V ow = h.valid ? loadv(h.gi, h.gj, mb + u + 2) : V{0, 0};
                    if (fwd && !(mb + u + 2 >= 0 && mb + u + 2 < nz)) ow.l = INFF;
                    q[0][u] = ow;
                    q[1][u] = (h.uxm || h.uxp) ? loadv(h.hxi, h.hxj, mb + u) : V{0, 0};
                    q[2][u] = (h.uym || h.uyp) ? loadv(h.hyi, h.hyj, mb + u) : V{0, 0};
                    fv[u] = h.valid ? ld(T, h.gi, h.gj, mb + u) : 0.0f;  // cell time (fwd)
                }
            };
            for (int t = 0; t < NT; t++) {
                Th &h = th[t];
                h.li = t / TJ;
                h.lj = t % TJ;
                h.valid = h.li < TIe && h.lj < TJe;
                h.pipj = h.valid ? (s1 > 0 ? h.li : TIe - 1 - h.li) + (s2 > 0 ? h.lj : TJe - 1 - h.lj) : 0;
#ifdef PLANT_NO_SKEW  // planted mistake (tests only): every column on the same level
                h.pipj = 0;
#endif
                const bool exm = h.lj == 0, exp_ = h.lj == TJe - 1, eym = h.li == 0, eyp = h.li == TIe - 1;
                h.gi = i0 + h.li;
                h.gj = j0 + h.lj;
                h.uxm = h.valid && exm && j0 > 0;
                h.uxp = h.valid && !h.uxm && exp_ && j0 + TJe < nc;
                h.uym = h.valid && eym && i0 > 0;
                h.uyp = h.valid && !h.uym && eyp && i0 + TIe < nr;
#ifdef PLANT_SWAP_EXCHANGE  // planted mistake (tests only): x- reads the wrong slot
                h.rxm = !h.valid ? t : exm ? INFSLOT : (s2 > 0 ? NT : 0) + t - 1;
#else
                h.rxm = !h.valid ? t : exm ? INFSLOT : (s2 > 0 ? 0 : NT) + t - 1;
#endif
                h.rxp = !h.valid ? t : exp_ ? INFSLOT : (s2 > 0 ? NT : 0) + t + 1;
                h.rym = !h.valid ? t : eym ? INFSLOT : (s1 > 0 ? 0 : NT) + t - TJ;
                h.ryp = !h.valid ? t : eyp ? INFSLOT : (s1 > 0 ? NT : 0) + t + TJ;
                h.hxi = h.gi;
                h.hxj = h.uxm ? j0 - 1 : j0 + TJe;
                h.hyi = h.uym ? i0 - 1 : i0 + TIe;
                h.hyj = h.gj;
                h.inter = h.valid && h.gi >= 1 && h.gi <= nr - 2 && h.gj >= 1 && h.gj <= nc - 2;
                if (fwd) {
                    h.plo = h.valid ? h.pipj : 0x3fffffff;
                    h.phi = h.pipj + nz;
                } else {
                    h.plo = h.inter ? h.pipj + 1 : 0x3fffffff;
                    h.phi = h.pipj + nz - 1;
                }
                h.psrc = -1;
                if (fwd && h.valid && src >= 0 && src % plane == (long)h.gi * nc + h.gj) {
                    const int sk = (int)(src / plane);
                    h.psrc = h.pipj + (s3 > 0 ? sk : nz - 1 - sk);
                }
                for (int k = 0; k < 3; k++) {
                    h.q[k].assign(L, V{0, 0});
                    h.nq[k].assign(L, V{0, 0});
                }
                h.fq.assign(L, 0);
                h.nfq.assign(L, 0);
                h.sval.assign(L, 0);
                h.chg.assign(L, 0);
                const int m0 = -h.pipj;
                const V none = V{fwd ? INFF : 0.0f, 0.0f};
                h.prev = none;
                h.cur = h.valid && m0 == 0 ? loadv(h.gi, h.gj, 0) : none;
                h.next = h.valid && m0 >= -1 ? loadv(h.gi, h.gj, m0 + 1) : none;
                load_batch(h, 0, h.q, h.fq);
            }
            auto batch = [&](int pb, bool swapq) {
                for (int t = 0; t < NT; t++) {  // stores of the previous batch, loads of the next
                    Th &h = th[t];
                    for (int u = 0; u < L; u++)
                        if (h.chg[u]) stv(h.gi, h.gj, pb - L - h.pipj + u, h.sval[u]);
                    if (!swapq) load_batch(h, pb + L, h.nq, h.nfq);
                    else load_batch(h, pb + L, h.q, h.fq);
                }
                for (int u = 0; u < L; u++) {
                    const int p = pb + u;
                    const int par = (u & 1) * PB;
                    for (int t = 0; t < NT; t++) {
                        pub[par + t] = th[t].prev;
                        pub[par + NT + t] = th[t].next;
                    }
                    for (int t = 0; t < NT; t++) {
                        Th &h = th[t];
                        std::vector<V> *q = swapq ? h.nq : h.q;
                        const float fq = swapq ? h.nfq[u] : h.fq[u];
                        const V pxm = pub[par + h.rxm], pxp = pub[par + h.rxp];
                        const V pym = pub[par + h.rym], pyp = pub[par + h.ryp];
                        bool c = false;
                        float nv = h.cur.l;
                        if (fwd) {
                            const float xa = fminf(h.uxm ? q[1][u].l : pxm.l, h.uxp ? q[1][u].l : pxp.l);
                            const float ya = fminf(h.uym ? q[2][u].l : pym.l, h.uyp ? q[2][u].l : pyp.l);
                            const float za = fminf(h.prev.l, h.next.l);
                            if (p >= h.plo && p < h.phi && p != h.psrc) {
                                const float qv = eik_uni(xa, ya, za, fq);
                                c = qv < h.cur.l;
                                nv = c ? qv : h.cur.l;
                            }
                        } else if (p >= h.plo && p < h.phi && !std::signbit(h.cur.t)) {
                            const V xm = h.uxm ? q[1][u] : pxm, xp = h.uxp ? q[1][u] : pxp;
                            const V ym = h.uym ? q[2][u] : pym, yp = h.uyp ? q[2][u] : pyp;
#ifndef PLANT_RING_DEPTH
                            const V zm = s3 > 0 ? h.prev : h.next, zp = s3 > 0 ? h.next : h.prev;
#else
                            const V zm = h.prev, zp = h.next;
#endif
                            const float nw = adj(h.cur.t, fabsf(xm.t), fabsf(xp.t), fabsf(ym.t),
                                                 fabsf(yp.t), fabsf(zm.t), fabsf(zp.t), xm.l, xp.l,
                                                 ym.l, yp.l, zm.l, zp.l);
                            c = nw != h.cur.l;
                            nv = nw;
                        }
                        h.chg[u] = c;
                        h.sval[u] = nv;
                        h.prev = V{nv, h.cur.t};
                        h.cur = h.next;
                        h.next = q[0][u];
                    }
                }
            };
            // batches until the one holding step pend-1, the queues alternating
            const int pend = nz + span;
            for (int pb = 0;; pb += 2 * L) {
                batch(pb, false);
                if (pb + L >= pend) break;
                batch(pb + L, true);
                if (pb + 2 * L >= pend) break;
            }
            const int pl = (pend - 1) / L * L;  // the last batch
            for (int t = 0; t < NT; t++) {
                Th &h = th[t];
                for (int u = 0; u < L; u++)
                    if (h.chg[u]) stv(h.gi, h.gj, pl - h.pipj + u, h.sval[u]);
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
            col(fwd, T.data(), B.data(), src, o, TI, TJ, L);
            for (long m = 0; m < N; m++)
                if (A[m] != B[m] && !(std::isnan(A[m]) && std::isnan(B[m]))) bad++;
            B = A;  // continue both from the same state
        }
    printf("orderings %d mismatches %ld\n", n, bad);
    return bad ? 1 : 0;
}
