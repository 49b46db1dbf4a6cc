/*
 * CPU oracle and reference-equivalent CPU baseline for the simulate path (float64, plain C).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ as the checker and by bench.py's cpu_baseline leg as the
 * timed CPU column. The product never links it.
 *
 * It restates the reference's per-event algorithm with no propagator reuse:
 *   bmc_matrix          src/bmc_jax/simulation/matrix.py:83-295
 *   the time step       src/bmc_jax/simulation/expm.py:94-112 (augmented [[A, C], [0, 0]] * dt)
 *   the scan            src/bmc_jax/simulation/simulate.py:22-132 (dt > 1e-9, spoiler, ADC reset)
 * The float64 exponential is the scaling-and-squaring Pade method of Higham (2005), degrees
 * 3/5/7/9/13 with theta_13 = 5.371920351148152 -- the float64 algorithm behind
 * jax.scipy.linalg.expm, which the reference calls at expm.py:28. The Jacobian is forward mode: the
 * product rule through every Pade product, the solve and each squaring (what jax.jacfwd computes),
 * with the generator's derivative taken by central differences of bmc_matrix (an independent route
 * from the GPU kernel's analytic derivatives).
 *
 * Packed parameter record per voxel (stride 13 + 5 * nc), the layout of include/bmc_hip.h:
 *   0 water_r1 1 water_r2 2 water_f 3 mt_k 4 mt_f 5 mt_dw 6 mt_r1 7 mt_r2 8 mt_lineshape
 *   9 b0 10 gamma 11 b0_inhom 12 rel_b1, then cest_k[nc] cest_f[nc] cest_dw[nc] cest_r1[nc] cest_r2[nc]
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAXN 21
#define MAXP 16

static int stride_of(int nc) { return 13 + 5 * nc; }

static double sl_sum(double delta, double t2) {
    /* matrix.py:34-38 */
    double s = 0.0;
    for (int i = 0; i <= 100; i++) {
        double u = i * 0.01;
        double t1 = fabs(3.0 * u * u - 1.0);
        if (t1 < 1e-6) t1 = 1e-6;
        double arg = delta * t2 / t1;
        s += sqrt(2.0 / M_PI) * t2 / t1 * exp(-2.0 * arg * arg);
    }
    return s * M_PI * 0.01;
}

static double sl_chs(double delta, double t2, double w0) {
    /* matrix.py:40-60 */
    double px[4] = {-300 - w0, -100 - w0, 100 + w0, 300 + w0};
    double py[4];
    for (int i = 0; i < 4; i++) py[i] = sl_sum(px[i], t2);
    double d0y = 30.0 * (py[1] - py[0]), d1y = 30.0 * (py[3] - py[2]);
    double c = fabs((delta - px[1] + 1) / (px[2] - px[1] + 1));
    double c2 = c * c, c3 = c2 * c;
    return (2 * c3 - 3 * c2 + 1) * py[1] + (-2 * c3 + 3 * c2) * py[2] + (c3 - 2 * c2 + c) * d0y + (c3 - c2) * d1y;
}

/* Augmented generator X = [[A, C], [0, 0]] * dt, n = N + 1, row-major n x n. */
static void build_x(const double* p, int nc, double amp, double ph, double fr, double dt, double* X) {
    int nt = 1 + nc, N = 3 * nt + 1, n = N + 1;
    const double *k = p + 13, *f = p + 13 + nc, *dw = p + 13 + 2 * nc, *r1c = p + 13 + 3 * nc, *r2c = p + 13 + 4 * nc;
    double r1w = p[0], r2w = p[1], kmt = p[3], fmt = p[4], dwmt = p[5], r1mt = p[6], r2mt = p[7];
    int ls = (int)p[8];
    double w1 = 2 * M_PI * amp * p[12], w1x = w1 * cos(ph), w1y = w1 * sin(ph);
    double larmor = p[10] * 1e6 * p[9];
    double dw0 = larmor * p[11] * 1e-6;
    double eff = 2 * M_PI * fr - dw0;
    double off_w = eff, off_mt = eff - dwmt * 1e-6 * larmor;
    double g = 0.0;
    if (fmt > 1e-9) {
        double t2 = 1.0 / r2mt;
        if (ls == 1) g = t2 / (1 + (off_mt * t2) * (off_mt * t2));
        else if (ls == 2) g = (fabs(off_mt) >= larmor * 1e-6) ? sl_sum(off_mt, t2) : sl_chs(off_mt, t2, larmor * 1e-6);
    }
    double wmt = (fmt > 1e-9) ? w1 * w1 * g : 0.0;
    double ks = 0.0;
    for (int i = 0; i < nc; i++) ks += k[i] * f[i];
    double kwmt = kmt * fmt;
    memset(X, 0, sizeof(double) * n * n);
#define XA(r, c) X[(r) * n + (c)]
    int xw = 0, yw = nt, zw = 2 * nt, zmt = 3 * nt;
    XA(xw, xw) = -r2w - ks; XA(xw, yw) = off_w; XA(xw, zw) = -w1y;
    XA(yw, yw) = -r2w - ks; XA(yw, xw) = -off_w; XA(yw, zw) = w1x;
    XA(zw, zw) = -r1w - ks - kwmt; XA(zw, yw) = -w1x; XA(zw, xw) = w1y;
    for (int i = 0; i < nc; i++) {
        int xc = 1 + i, yc = nt + 1 + i, zc = 2 * nt + 1 + i;
        double off_c = dw[i] * 1e-6 * larmor - eff;
        XA(xc, xc) = -r2c[i] - k[i]; XA(xc, yc) = -off_c; XA(xc, zc) = -w1y;
        XA(yc, yc) = -r2c[i] - k[i]; XA(yc, xc) = off_c; XA(yc, zc) = w1x;
        XA(zc, zc) = -r1c[i] - k[i]; XA(zc, xc) = w1y; XA(zc, yc) = -w1x;
        XA(xw, xc) = k[i]; XA(yw, yc) = k[i]; XA(zw, zc) = k[i];
        XA(xc, xw) = k[i] * f[i]; XA(yc, yw) = k[i] * f[i]; XA(zc, zw) = k[i] * f[i];
        XA(zc, N) = r1c[i] * f[i];
    }
    XA(zmt, zmt) = -r1mt - kmt - wmt; XA(zmt, zw) = kwmt; XA(zw, zmt) = kmt;
    XA(zw, N) = r1w; XA(zmt, N) = r1mt * fmt;
    for (int i = 0; i < n * n; i++) X[i] *= dt;
#undef XA
}

static void mm(int n, const double* A, const double* B, double* C) {
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double s = 0.0;
            for (int k = 0; k < n; k++) s += A[i * n + k] * B[k * n + j];
            C[i * n + j] = s;
        }
}

/* C += A B */
static void mm_acc(int n, const double* A, const double* B, double* C) {
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double s = 0.0;
            for (int k = 0; k < n; k++) s += A[i * n + k] * B[k * n + j];
            C[i * n + j] += s;
        }
}

/* LU with partial pivoting of Q (in place), then solve for nrhs n x n right-hand sides in place. */
static void lu(int n, double* Q, int* piv) {
    for (int c = 0; c < n; c++) {
        int pr = c;
        double best = fabs(Q[c * n + c]);
        for (int r = c + 1; r < n; r++)
            if (fabs(Q[r * n + c]) > best) { best = fabs(Q[r * n + c]); pr = r; }
        piv[c] = pr;
        if (pr != c)
            for (int j = 0; j < n; j++) { double t = Q[c * n + j]; Q[c * n + j] = Q[pr * n + j]; Q[pr * n + j] = t; }
        for (int r = c + 1; r < n; r++) {
            double m = Q[r * n + c] / Q[c * n + c];
            Q[r * n + c] = m;
            for (int j = c + 1; j < n; j++) Q[r * n + j] -= m * Q[c * n + j];
        }
    }
}

static void lu_solve(int n, const double* LU, const int* piv, double* B) {
    /* all row interchanges first (the multipliers are stored in final row order), then unit-L */
    for (int c = 0; c < n; c++)
        if (piv[c] != c)
            for (int j = 0; j < n; j++) { double t = B[c * n + j]; B[c * n + j] = B[piv[c] * n + j]; B[piv[c] * n + j] = t; }
    for (int c = 0; c < n; c++) {
        for (int r = c + 1; r < n; r++) {
            double m = LU[r * n + c];
            for (int j = 0; j < n; j++) B[r * n + j] -= m * B[c * n + j];
        }
    }
    for (int c = n - 1; c >= 0; c--) {
        for (int j = 0; j < n; j++) B[c * n + j] /= LU[c * n + c];
        for (int r = 0; r < c; r++) {
            double m = LU[r * n + c];
            for (int j = 0; j < n; j++) B[r * n + j] -= m * B[c * n + j];
        }
    }
}

static const double B13[14] = {64764752532480000., 32382376266240000., 7771770303897600.,
    1187353796428800., 129060195264000., 10559470521600., 670442572800., 33522128640.,
    1323241920., 40840800., 960960., 16380., 182., 1.};

/*
 * E = expm(X) by Higham (2005) Pade 3/5/7/9/13 scaling and squaring, with P forward-mode tangents
 * dE[p] = L(X, dX[p]). Returns the squaring count (the caller applies the reference NaN rule).
 */
static int expm_dual(int n, const double* X, int P, const double* dX, double* E, double* dE, double* ws) {
    static const double th[5] = {1.495585217958292e-2, 2.539398330063230e-1, 9.504178996162932e-1,
                                 2.097847961257068e0, 5.371920351148152e0};
    static const double b3[4] = {120., 60., 12., 1.};
    static const double b5[6] = {30240., 15120., 3360., 420., 30., 1.};
    static const double b7[8] = {17297280., 8648640., 1995840., 277200., 25200., 1512., 56., 1.};
    static const double b9[10] = {17643225600., 8821612800., 2075673600., 302702400., 30270240.,
                                  2162160., 110880., 3960., 90., 1.};
    int nn = n * n;
    double l1 = 0.0;
    for (int j = 0; j < n; j++) {
        double s = 0.0;
        for (int i = 0; i < n; i++) s += fabs(X[i * n + j]);
        if (s > l1) l1 = s;
    }
    int m, s = 0;
    if (l1 <= th[0]) m = 3; else if (l1 <= th[1]) m = 5; else if (l1 <= th[2]) m = 7;
    else if (l1 <= th[3]) m = 9; else { m = 13; s = (int)ceil(log2(l1 / th[4])); if (s < 0) s = 0; }
    double sc = ldexp(1.0, -s);
    double A[MAXN * MAXN], A2[MAXN * MAXN], A4[MAXN * MAXN], A6[MAXN * MAXN], T[MAXN * MAXN];
    double U[MAXN * MAXN], V[MAXN * MAXN], W[MAXN * MAXN];
    /* tangent arrays live in the caller's workspace: 9 blocks of MAXP x MAXN^2 */
    typedef double Blk[MAXN * MAXN];
    Blk* dA = (Blk*)(ws);
    Blk* dA2 = dA + MAXP;
    Blk* dA4 = dA2 + MAXP;
    Blk* dA6 = dA4 + MAXP;
    Blk* dU = dA6 + MAXP;
    Blk* dV = dU + MAXP;
    Blk* dW = dV + MAXP;
    Blk* dR = dW + MAXP;
    Blk* dA8 = dR + MAXP;
    for (int i = 0; i < nn; i++) A[i] = X[i] * sc;
    for (int p = 0; p < P; p++) for (int i = 0; i < nn; i++) dA[p][i] = dX[p * nn + i] * sc;
    mm(n, A, A, A2);
    for (int p = 0; p < P; p++) { mm(n, dA[p], A, dA2[p]); mm_acc(n, A, dA[p], dA2[p]); }
    if (m >= 5) {
        mm(n, A2, A2, A4);
        for (int p = 0; p < P; p++) { mm(n, dA2[p], A2, dA4[p]); mm_acc(n, A2, dA2[p], dA4[p]); }
    }
    if (m >= 7) {
        mm(n, A4, A2, A6);
        for (int p = 0; p < P; p++) { mm(n, dA4[p], A2, dA6[p]); mm_acc(n, A4, dA2[p], dA6[p]); }
    }
    /* W = odd polynomial without the leading A; V = even polynomial. */
    const double* b = (m == 3) ? b3 : (m == 5) ? b5 : (m == 7) ? b7 : (m == 9) ? b9 : B13;
    memset(W, 0, sizeof(double) * nn); memset(V, 0, sizeof(double) * nn);
    for (int p = 0; p < P; p++) { memset(dW[p], 0, sizeof(double) * nn); memset(dV[p], 0, sizeof(double) * nn); }
    if (m < 13) {
        const double* pw[5] = {0, A2, A4, A6, 0};
        double A8[MAXN * MAXN];
        if (m == 9) {
            mm(n, A4, A4, A8);
            for (int p = 0; p < P; p++) { mm(n, dA4[p], A4, dA8[p]); mm_acc(n, A4, dA4[p], dA8[p]); }
            pw[4] = A8;
        }
        for (int i = 0; i < n; i++) { W[i * n + i] += b[1]; V[i * n + i] += b[0]; }
        for (int q = 1; 2 * q <= m - 1; q++) {
            const double* Pq = pw[q];
            for (int i = 0; i < nn; i++) { W[i] += b[2 * q + 1] * Pq[i]; V[i] += b[2 * q] * Pq[i]; }
            for (int p = 0; p < P; p++) {
                const double* dPq = (q == 1) ? dA2[p] : (q == 2) ? dA4[p] : (q == 3) ? dA6[p] : dA8[p];
                for (int i = 0; i < nn; i++) { dW[p][i] += b[2 * q + 1] * dPq[i]; dV[p][i] += b[2 * q] * dPq[i]; }
            }
        }
    } else {
        /* W = A6 (b13 A6 + b11 A4 + b9 A2) + b7 A6 + b5 A4 + b3 A2 + b1 I ; V likewise (even). */
        for (int i = 0; i < nn; i++) T[i] = b[13] * A6[i] + b[11] * A4[i] + b[9] * A2[i];
        mm(n, A6, T, W);
        for (int i = 0; i < nn; i++) W[i] += b[7] * A6[i] + b[5] * A4[i] + b[3] * A2[i];
        for (int i = 0; i < n; i++) W[i * n + i] += b[1];
        for (int p = 0; p < P; p++) {
            double dT[MAXN * MAXN];
            for (int i = 0; i < nn; i++) dT[i] = b[13] * dA6[p][i] + b[11] * dA4[p][i] + b[9] * dA2[p][i];
            mm(n, dA6[p], T, dW[p]); mm_acc(n, A6, dT, dW[p]);
            for (int i = 0; i < nn; i++) dW[p][i] += b[7] * dA6[p][i] + b[5] * dA4[p][i] + b[3] * dA2[p][i];
        }
        for (int i = 0; i < nn; i++) T[i] = b[12] * A6[i] + b[10] * A4[i] + b[8] * A2[i];
        mm(n, A6, T, V);
        for (int i = 0; i < nn; i++) V[i] += b[6] * A6[i] + b[4] * A4[i] + b[2] * A2[i];
        for (int i = 0; i < n; i++) V[i * n + i] += b[0];
        for (int p = 0; p < P; p++) {
            double dT[MAXN * MAXN];
            for (int i = 0; i < nn; i++) dT[i] = b[12] * dA6[p][i] + b[10] * dA4[p][i] + b[8] * dA2[p][i];
            mm(n, dA6[p], T, dV[p]); mm_acc(n, A6, dT, dV[p]);
            for (int i = 0; i < nn; i++) dV[p][i] += b[6] * dA6[p][i] + b[4] * dA4[p][i] + b[2] * dA2[p][i];
        }
    }
    mm(n, A, W, U);
    for (int p = 0; p < P; p++) { mm(n, dA[p], W, dU[p]); mm_acc(n, A, dW[p], dU[p]); }
    /* R = (V - U)^-1 (V + U); dR = (V - U)^-1 (dV + dU - (dV - dU) R) */
    double Q[MAXN * MAXN], R[MAXN * MAXN];
    int piv[MAXN];
    for (int i = 0; i < nn; i++) { Q[i] = V[i] - U[i]; R[i] = V[i] + U[i]; }
    lu(n, Q, piv);
    lu_solve(n, Q, piv, R);
    for (int p = 0; p < P; p++) {
        double dQ[MAXN * MAXN];
        for (int i = 0; i < nn; i++) dQ[i] = dV[p][i] - dU[p][i];
        mm(n, dQ, R, T);
        for (int i = 0; i < nn; i++) dR[p][i] = dV[p][i] + dU[p][i] - T[i];
        lu_solve(n, Q, piv, dR[p]);
    }
    for (int k = 0; k < s; k++) {
        for (int p = 0; p < P; p++) {
            mm(n, dR[p], R, T); mm_acc(n, R, dR[p], T);
            memcpy(dR[p], T, sizeof(double) * nn);
        }
        mm(n, R, R, T);
        memcpy(R, T, sizeof(double) * nn);
    }
    memcpy(E, R, sizeof(double) * nn);
    for (int p = 0; p < P; p++) memcpy(dE + p * nn, dR[p], sizeof(double) * nn);
    return s;
}

static int ref32_squarings(const double* X, int n) {
    /* reference float32 rule, expm.py:73-74: floor(log2(l1 / 3.925724783138660)), used for NaN > 16 */
    float l1 = 0.f;
    for (int j = 0; j < n; j++) {
        float s = 0.f;
        for (int i = 0; i < n; i++) s += fabsf((float)X[i * n + j]);
        if (s > l1) l1 = s;
    }
    if (l1 <= 0.f) return 0;
    double v = floor(log2((double)l1 / 3.925724783138660));
    return v < 0 ? 0 : (int)v;
}

/*
 * One offset row of one voxel. Returns last_M (N) and its P tangents w.r.t. packed parameter
 * slots (physical units). Generator derivatives by central differences with relative step rel_h.
 */
static void sim_row(const double* p, int nc, int T, const int* typ, const double* dt, const double* amp,
                    const double* ph, const double* fr, const double* m0, const double* dm0, int P,
                    const int* slot, double rel_h, int nan_rule, double* out_m, double* out_dm) {
    int nt = 1 + nc, N = 3 * nt + 1, n = N + 1, nn = n * n, st = stride_of(nc);
    double M[MAXN], last[MAXN], dM[MAXP][MAXN], dlast[MAXP][MAXN];
    double X[MAXN * MAXN], E[MAXN * MAXN];
    double* heap = (double*)malloc(sizeof(double) * (size_t)MAXP * MAXN * MAXN * 11);
    double* dX = heap;
    double* dE = heap + (size_t)MAXP * MAXN * MAXN;
    double* ws = heap + (size_t)2 * MAXP * MAXN * MAXN;
    double Xp[MAXN * MAXN], Xm[MAXN * MAXN], q[13 + 5 * 8];
    for (int i = 0; i < N; i++) { M[i] = m0[i]; last[i] = M[i]; }
    for (int k = 0; k < P; k++) for (int i = 0; i < N; i++) { dM[k][i] = dm0 ? dm0[k * N + i] : 0.0; dlast[k][i] = dM[k][i]; }
    for (int e = 0; e < T; e++) {
        if (dt[e] > 1e-9) {
            build_x(p, nc, amp[e], ph[e], fr[e], dt[e], X);
            for (int k = 0; k < P; k++) {
                memcpy(q, p, sizeof(double) * st);
                double h = rel_h * fmax(fabs(p[slot[k]]), 1e-3);
                q[slot[k]] = p[slot[k]] + h; build_x(q, nc, amp[e], ph[e], fr[e], dt[e], Xp);
                q[slot[k]] = p[slot[k]] - h; build_x(q, nc, amp[e], ph[e], fr[e], dt[e], Xm);
                for (int i = 0; i < nn; i++) dX[k * nn + i] = (Xp[i] - Xm[i]) / (2 * h);
            }
            if (nan_rule && ref32_squarings(X, n) > 16) {
                for (int i = 0; i < N; i++) M[i] = NAN;
                for (int k = 0; k < P; k++) for (int i = 0; i < N; i++) dM[k][i] = NAN;
            } else {
                expm_dual(n, X, P, dX, E, dE, ws);
                double Mn[MAXN], dMn[MAXP][MAXN];
                for (int i = 0; i < N; i++) {
                    double s = E[i * n + N];
                    for (int j = 0; j < N; j++) s += E[i * n + j] * M[j];
                    Mn[i] = s;
                }
                for (int k = 0; k < P; k++)
                    for (int i = 0; i < N; i++) {
                        double s = dE[k * nn + i * n + N];
                        for (int j = 0; j < N; j++) s += dE[k * nn + i * n + j] * M[j] + E[i * n + j] * dM[k][j];
                        dMn[k][i] = s;
                    }
                memcpy(M, Mn, sizeof(double) * N);
                for (int k = 0; k < P; k++) memcpy(dM[k], dMn[k], sizeof(double) * N);
            }
        }
        if (typ[e] == 2 || typ[e] == 4) {
            for (int i = 0; i < 2 * nt; i++) M[i] = 0.0;
            for (int k = 0; k < P; k++) for (int i = 0; i < 2 * nt; i++) dM[k][i] = 0.0;
        }
        memcpy(last, M, sizeof(double) * N);
        for (int k = 0; k < P; k++) memcpy(dlast[k], dM[k], sizeof(double) * N);
        if (typ[e] == 3) {
            for (int i = 0; i < N; i++) M[i] = m0[i];
            for (int k = 0; k < P; k++) for (int i = 0; i < N; i++) dM[k][i] = dm0 ? dm0[k * N + i] : 0.0;
        }
    }
    memcpy(out_m, last, sizeof(double) * N);
    for (int k = 0; k < P; k++) memcpy(out_dm + k * N, dlast[k], sizeof(double) * N);
    free(heap);
}

/*
 * Simulate V voxels x O offset rows x T events. params [V, 13 + 5 nc]; events [O, T]; m0 [V, N];
 * dm0 [V, P, N] or NULL. Outputs: m_out [V, O, N] (may be NULL), dm_out [V, O, P, N] (may be NULL).
 * Returns 0, or -1 on bad sizes.
 */
int oracle_simulate(int V, int nc, int O, int T, const double* params, const int* typ, const double* dt,
                    const double* amp, const double* ph, const double* fr, const double* m0,
                    const double* dm0, int P, const int* slot, double rel_h, int nan_rule, int n_threads,
                    double* m_out, double* dm_out) {
    int N = 3 * (1 + nc) + 1;
    if (nc < 0 || nc > 8 || N + 1 > MAXN || P < 0 || P > MAXP) return -1;
    int st = stride_of(nc);
    long total = (long)V * O;
#pragma omp parallel for schedule(dynamic, 1) num_threads(n_threads > 0 ? n_threads : 1)
    for (long t = 0; t < total; t++) {
        int v = (int)(t / O), o = (int)(t % O);
        double mo[MAXN], dmo[MAXP * MAXN];
        sim_row(params + (long)v * st, nc, T, typ + (long)o * T, dt + (long)o * T, amp + (long)o * T,
                ph + (long)o * T, fr + (long)o * T, m0 + (long)v * N, dm0 ? dm0 + (long)v * P * N : NULL,
                P, slot, rel_h, nan_rule, mo, dmo);
        if (m_out) memcpy(m_out + t * N, mo, sizeof(double) * N);
        if (dm_out) memcpy(dm_out + t * P * N, dmo, sizeof(double) * P * N);
    }
    return 0;
}

/* Batched float64 expm with tangents (for checking the oracle itself against scipy). */
int oracle_expm(int B, int n, const double* X, int P, const double* dX, double* E, double* dE) {
    if (n > MAXN || P > MAXP) return -1;
    double* ws = (double*)malloc(sizeof(double) * (size_t)MAXP * MAXN * MAXN * 9);
    for (int b = 0; b < B; b++)
        expm_dual(n, X + (long)b * n * n, (dX && dE) ? P : 0, dX ? dX + (long)b * P * n * n : NULL, E + (long)b * n * n,
                  dE ? dE + (long)b * P * n * n : NULL, ws);
    free(ws);
    return 0;
}
