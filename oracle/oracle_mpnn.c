/*
 * oracle_mpnn.c -- FP64 CPU restatement of the learned warm-start network.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  The reference has no learned
 * component: its LOBPCG start block is N(0,1) plus the sqrt-degree hint
 * (src/solvers/lobpcg.rs:243-252) and this network predicts that block
 * instead.  The functions below define the network's mathematics; the device
 * kernels (bf16 matrix cores, FP32 accumulation) are checked against them, and
 * they are pinned by central finite differences (tests/test_oracle_mpnn.py).
 *
 * Graph operator: A_hat = D^-1/2 W D^-1/2 without its diagonal, i.e. I - L,
 * given as CSR (a_indptr, a_indices, a_data) of the off-diagonal entries.
 *
 * Node features F [n x 16]:
 *   0 log1p(deg)            1 log1p(#neighbours)
 *   2..5 mean / max / min / std of the row's A_hat entries
 *   6..13 random-walk probes A_hat^t p_0 recorded at walk lengths
 *         t = 1, 2, 3, 4, 6, 8, 12, 16; p_0 = N(0,1) draws (Philox stream 9,
 *         index = node id) with the trivial direction D^{1/2} 1 projected
 *         out; each probe scaled to unit RMS
 *   14 sqrt(deg) / max sqrt(deg)   15 1
 *
 * Network (hidden 64, tanh), parameters packed in this order:
 *   layer l = 1..3: Ws_l [d_in x 64], Wn_l [d_in x 64], b_l [64]  (d_in = 16, 64, 64)
 *   output: Wo [64 x 8]
 *   H_0 = F;  A_l = A_hat H_{l-1};  H_l = tanh(H_{l-1} Ws_l + A_l Wn_l + b_l);  X = H_3 Wo
 *
 * The 8 output columns are a subspace, not the eigenvectors themselves: the
 * warm start takes the Rayleigh-Ritz vectors of [D^{1/2} 1 | X].  The loss is
 * how much of the target V [n x 2] (the two non-trivial eigenvectors the
 * solver returns) lies outside span(X), with a ridge that keeps the Gram
 * matrix invertible:
 *   G = X^T X,  B = X^T V,  delta = RIDGE tr(G) / 8,  C = (G + delta I)^-1 B
 *   loss = tr(V^T V) - tr(B^T C)     (sum_j ||v_j - P_X v_j||^2 as RIDGE -> 0)
 *   dloss/dX = -2 V C^T + 2 X (C C^T + (RIDGE / 4) tr(C^T C) I)
 * invariant to the scale of X and to the sign and rotation of V.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define HID 64
#define NF 16
#define NOUT 8
#define RIDGE 1e-3

int64_t or_mpnn_param_count(void) {
    return (int64_t)(NF * HID * 2 + HID) + 2 * (int64_t)(HID * HID * 2 + HID) + HID * NOUT;
}

/* offsets of each block in the packed parameter vector */
static void param_layout(int64_t off[10]) {
    int64_t o = 0;
    int din = NF;
    for (int l = 0; l < 3; ++l) {
        off[3 * l] = o; o += (int64_t)din * HID;        /* Ws */
        off[3 * l + 1] = o; o += (int64_t)din * HID;    /* Wn */
        off[3 * l + 2] = o; o += HID;                   /* b  */
        din = HID;
    }
    off[9] = o;                                         /* Wo */
}

static void spmm_offdiag(int64_t n, const int64_t* ip, const int32_t* ix, const double* a,
                         const double* x, int k, double* y) {
    for (int64_t i = 0; i < n; ++i) {
        double* yi = y + i * k;
        for (int c = 0; c < k; ++c) yi[c] = 0.0;
        for (int64_t p = ip[i]; p < ip[i + 1]; ++p) {
            const double* xj = x + (int64_t)ix[p] * k;
            for (int c = 0; c < k; ++c) yi[c] += a[p] * xj[c];
        }
    }
}

void or_mpnn_features(int64_t n, const int64_t* ip, const int32_t* ix, const double* a,
                      const double* deg, uint64_t seed, double* f) {
    double sqmax = 0.0, sqn2 = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double sq = sqrt(deg[i] > 0.0 ? deg[i] : 0.0);
        if (sq > sqmax) sqmax = sq;
        sqn2 += sq * sq;
    }
    for (int64_t i = 0; i < n; ++i) {
        double* fi = f + i * NF;
        const int64_t b = ip[i], e = ip[i + 1];
        const double cnt = (double)(e - b);
        double sum = 0.0, mx = 0.0, mn = 0.0, ss = 0.0;
        for (int64_t p = b; p < e; ++p) {
            sum += a[p];
            if (p == b || a[p] > mx) mx = a[p];
            if (p == b || a[p] < mn) mn = a[p];
        }
        const double mean = cnt > 0 ? sum / cnt : 0.0;
        for (int64_t p = b; p < e; ++p) ss += (a[p] - mean) * (a[p] - mean);
        fi[0] = log1p(deg[i] > 0.0 ? deg[i] : 0.0);
        fi[1] = log1p(cnt);
        fi[2] = mean;
        fi[3] = mx;
        fi[4] = mn;
        fi[5] = cnt > 0 ? sqrt(ss / cnt) : 0.0;
        fi[14] = sqmax > 0.0 ? sqrt(deg[i] > 0.0 ? deg[i] : 0.0) / sqmax : 0.0;
        fi[15] = 1.0;
    }
    /* probes */
    double* p = (double*)malloc(sizeof(double) * (size_t)n);
    double* q = (double*)malloc(sizeof(double) * (size_t)n);
    double dot = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        p[i] = or_normal(seed, 9u, (uint64_t)i);
        dot += p[i] * sqrt(deg[i] > 0.0 ? deg[i] : 0.0);
    }
    if (sqn2 > 0.0)
        for (int64_t i = 0; i < n; ++i) p[i] -= dot / sqn2 * sqrt(deg[i] > 0.0 ? deg[i] : 0.0);
    static const int steps[8] = {1, 2, 3, 4, 6, 8, 12, 16};
    int done = 0;
    for (int t = 0; t < 8; ++t) {
        for (; done < steps[t]; ++done) {
            spmm_offdiag(n, ip, ix, a, p, 1, q);
            double* tmp = p; p = q; q = tmp;
        }
        double s2 = 0.0;
        for (int64_t i = 0; i < n; ++i) s2 += p[i] * p[i];
        const double rms = sqrt(s2 / (double)n);
        for (int64_t i = 0; i < n; ++i) f[i * NF + 6 + t] = rms > 0.0 ? p[i] / rms : 0.0;
    }
    free(p);
    free(q);
}

/* Forward pass; h [4 x n x HID] (h[0] holds F in its first 16 columns with
 * row stride NF), av [3 x n x HID] the A_hat products, x [n x 2]. */
void or_mpnn_forward(int64_t n, const int64_t* ip, const int32_t* ix, const double* a,
                     const double* f, const double* params, double* h, double* av, double* x) {
    int64_t off[10];
    param_layout(off);
    const double* hin = f;
    int din = NF;
    for (int l = 0; l < 3; ++l) {
        double* al = av + (int64_t)l * n * HID;
        double* hl = h + (int64_t)(l + 1) * n * HID;
        spmm_offdiag(n, ip, ix, a, hin, din, al);
        const double* ws = params + off[3 * l];
        const double* wn = params + off[3 * l + 1];
        const double* bb = params + off[3 * l + 2];
        for (int64_t i = 0; i < n; ++i)
            for (int c = 0; c < HID; ++c) {
                double z = bb[c];
                for (int k = 0; k < din; ++k)
                    z += hin[i * din + k] * ws[k * HID + c] + al[i * din + k] * wn[k * HID + c];
                hl[i * HID + c] = tanh(z);
            }
        hin = hl;
        din = HID;
    }
    const double* wo = params + off[9];
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < NOUT; ++c) {
            double s = 0.0;
            for (int k = 0; k < HID; ++k) s += hin[i * HID + k] * wo[k * NOUT + c];
            x[i * NOUT + c] = s;
        }
    if (h) memcpy(h, f, sizeof(double) * (size_t)n * NF);
}

/* Cholesky solve of the SPD system a [m x m] c = b [m x r] in place of b */
static int chol_solve(int m, double* a, double* b, int r) {
    for (int j = 0; j < m; ++j) {
        double d = a[j * m + j];
        for (int k = 0; k < j; ++k) d -= a[j * m + k] * a[j * m + k];
        if (!(d > 0.0)) return -1;
        d = sqrt(d);
        a[j * m + j] = d;
        for (int i = j + 1; i < m; ++i) {
            double e = a[i * m + j];
            for (int k = 0; k < j; ++k) e -= a[i * m + k] * a[j * m + k];
            a[i * m + j] = e / d;
        }
    }
    for (int c = 0; c < r; ++c) {
        for (int i = 0; i < m; ++i) {
            double e = b[i * r + c];
            for (int k = 0; k < i; ++k) e -= a[i * m + k] * b[k * r + c];
            b[i * r + c] = e / a[i * m + i];
        }
        for (int i = m - 1; i >= 0; --i) {
            double e = b[i * r + c];
            for (int k = i + 1; k < m; ++k) e -= a[k * m + i] * b[k * r + c];
            b[i * r + c] = e / a[i * m + i];
        }
    }
    return 0;
}

double or_mpnn_loss(int64_t n, const double* x, const double* v, double* dx) {
    double g[NOUT * NOUT] = {0}, b[NOUT * 2] = {0}, vv = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double* xi = x + i * NOUT;
        const double* vi = v + i * 2;
        for (int p = 0; p < NOUT; ++p) {
            for (int q = 0; q < NOUT; ++q) g[p * NOUT + q] += xi[p] * xi[q];
            b[p * 2] += xi[p] * vi[0];
            b[p * 2 + 1] += xi[p] * vi[1];
        }
        vv += vi[0] * vi[0] + vi[1] * vi[1];
    }
    double tr = 0.0;
    for (int p = 0; p < NOUT; ++p) tr += g[p * NOUT + p];
    const double delta = RIDGE * tr / NOUT;
    double gd[NOUT * NOUT], c[NOUT * 2];
    memcpy(gd, g, sizeof gd);
    for (int p = 0; p < NOUT; ++p) gd[p * NOUT + p] += delta;
    memcpy(c, b, sizeof c);
    if (chol_solve(NOUT, gd, c, 2) != 0) return NAN;
    double btc = 0.0, ctc = 0.0;
    for (int e = 0; e < NOUT * 2; ++e) { btc += b[e] * c[e]; ctc += c[e] * c[e]; }
    if (dx) {
        double pm[NOUT * NOUT], qm[2 * NOUT];
        for (int p = 0; p < NOUT; ++p)
            for (int q = 0; q < NOUT; ++q)
                pm[p * NOUT + q] = 2.0 * (c[p * 2] * c[q * 2] + c[p * 2 + 1] * c[q * 2 + 1]) +
                                   (p == q ? 2.0 * RIDGE / NOUT * ctc : 0.0);
        for (int t = 0; t < 2; ++t)
            for (int q = 0; q < NOUT; ++q) qm[t * NOUT + q] = -2.0 * c[q * 2 + t];
        for (int64_t i = 0; i < n; ++i)
            for (int q = 0; q < NOUT; ++q) {
                double d = v[i * 2] * qm[q] + v[i * 2 + 1] * qm[NOUT + q];
                for (int p = 0; p < NOUT; ++p) d += x[i * NOUT + p] * pm[p * NOUT + q];
                dx[i * NOUT + q] = d;
            }
    }
    return vv - btc;
}

/* Backward pass for a loss gradient dx [n x 8]; grad (packed like params). */
void or_mpnn_backward(int64_t n, const int64_t* ip, const int32_t* ix, const double* a,
                      const double* f, const double* params, const double* h,
                      const double* av, const double* dx, double* grad) {
    int64_t off[10];
    param_layout(off);
    memset(grad, 0, sizeof(double) * (size_t)or_mpnn_param_count());
    const double* h3 = h + 3 * n * HID;
    const double* wo = params + off[9];
    double* dh = (double*)malloc(sizeof(double) * (size_t)n * HID);
    double* dz = (double*)malloc(sizeof(double) * (size_t)n * HID);
    double* u = (double*)malloc(sizeof(double) * (size_t)n * HID);
    double* t = (double*)malloc(sizeof(double) * (size_t)n * HID);
    for (int k = 0; k < HID; ++k)
        for (int c = 0; c < NOUT; ++c) {
            double s = 0.0;
            for (int64_t i = 0; i < n; ++i) s += h3[i * HID + k] * dx[i * NOUT + c];
            grad[off[9] + k * NOUT + c] = s;
        }
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < HID; ++k) {
            double s = 0.0;
            for (int c = 0; c < NOUT; ++c) s += dx[i * NOUT + c] * wo[k * NOUT + c];
            dh[i * HID + k] = s;
        }
    for (int l = 2; l >= 0; --l) {
        const int din = l == 0 ? NF : HID;
        const double* hl = h + (int64_t)(l + 1) * n * HID;
        const double* hin = l == 0 ? f : h + (int64_t)l * n * HID;
        const double* al = av + (int64_t)l * n * HID;
        const double* ws = params + off[3 * l];
        const double* wn = params + off[3 * l + 1];
        for (int64_t e = 0; e < n * HID; ++e) dz[e] = dh[e] * (1.0 - hl[e] * hl[e]);
        for (int k = 0; k < din; ++k)
            for (int c = 0; c < HID; ++c) {
                double s1 = 0.0, s2 = 0.0;
                for (int64_t i = 0; i < n; ++i) {
                    s1 += hin[i * din + k] * dz[i * HID + c];
                    s2 += al[i * din + k] * dz[i * HID + c];
                }
                grad[off[3 * l] + k * HID + c] = s1;
                grad[off[3 * l + 1] + k * HID + c] = s2;
            }
        for (int c = 0; c < HID; ++c) {
            double s = 0.0;
            for (int64_t i = 0; i < n; ++i) s += dz[i * HID + c];
            grad[off[3 * l + 2] + c] = s;
        }
        if (l == 0) break;
        /* dh_{l-1} = dz Ws^T + A_hat (dz Wn^T) */
        for (int64_t i = 0; i < n; ++i)
            for (int k = 0; k < din; ++k) {
                double s = 0.0;
                for (int c = 0; c < HID; ++c) s += dz[i * HID + c] * wn[k * HID + c];
                u[i * din + k] = s;
            }
        spmm_offdiag(n, ip, ix, a, u, din, t);
        for (int64_t i = 0; i < n; ++i)
            for (int k = 0; k < din; ++k) {
                double s = t[i * din + k];
                for (int c = 0; c < HID; ++c) s += dz[i * HID + c] * ws[k * HID + c];
                dh[i * din + k] = s;
            }
    }
    free(dh); free(dz); free(u); free(t);
}
