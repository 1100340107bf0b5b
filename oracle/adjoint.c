/*
 * adjoint.c -- the differentiable relaxation of tn_adj_* (include/trussnet_amd.h)
 * in double precision on the CPU, forward and reverse: the FP64 gradient the
 * device adjoint is checked against. TEST INFRASTRUCTURE ONLY.
 *
 * The model (no counterpart in the reference): sizes trilinear in a grid theta, the
 * cell fractions through smoothstep f^2 (3 - 2 f) (C1 across cell faces)
 * (positions clamped to the grid), bars i -> j of the fixed lists with
 * l0 = c (h_i + h_j) / 2 (c = fsurf if neither end is interior, else fscale),
 * u = l0 - l, F_i = sum sp(u) e / l with sp(u) = log(1 + exp(beta u)) / beta,
 * D_i = 1 + sum s(beta u), x_i += dt F_i / D_i for interior nodes; after `iters`
 * steps L = (1/n) sum sp^2 / 2 + wq (1/ntet) sum (1 - q), q the Joe-Liu quality
 * 12 (3 Vs)^(2/3) / sum l^2 with Vs = sqrt(V^2 + eps_v^2).
 *
 * The reverse pass here is written the plain way -- every state kept, each bar's
 * contribution pushed to both of its ends -- so it shares no structure with the
 * device's checkpointed, gather-only adjoint. It is pinned by central finite
 * differences of the forward (tests/test_oracle_adjoint.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct {
    int i0[3];
    double f[3], gs[3];
} Tri;

static void tri_at(const tn_adj_problem* p, const double* x, Tri* t) {
    const int g[3] = { p->gx, p->gy, p->gz };
    for (int a = 0; a < 3; ++a) {
        double u = (x[a] - p->origin[a]) / p->spacing[a];
        t->gs[a] = 1.0 / p->spacing[a];
        if (!(u > 0.0)) {
            u = 0.0;
            t->gs[a] = 0.0;
        } else if (u >= g[a] - 1) {
            u = g[a] - 1;
            t->gs[a] = 0.0;
        }
        int i = (int)floor(u);
        if (i > g[a] - 2) {
            i = g[a] - 2;
        }
        t->i0[a] = i;
        /* the cell fraction through smoothstep: the interpolant is C1 across cell faces */
        const double f = u - i;
        t->f[a] = f * f * (3.0 - 2.0 * f);
        t->gs[a] *= 6.0 * f * (1.0 - f);
    }
}

static long tri_cell(const tn_adj_problem* p, const Tri* t, int c) {
    return (t->i0[0] + (c & 1)) + (long)p->gx * ((t->i0[1] + ((c >> 1) & 1)) + (long)p->gy * (t->i0[2] + (c >> 2)));
}

static double tri_w(const Tri* t, int c, int skip) {
    double w = 1.0;
    for (int a = 0; a < 3; ++a) {
        if (a == skip) {
            continue;
        }
        w *= ((c >> a) & 1) ? t->f[a] : 1.0 - t->f[a];
    }
    return w;
}

static double size_at(const tn_adj_problem* p, const double* theta, const double* x) {
    Tri t;
    tri_at(p, x, &t);
    double h = 0.0;
    for (int c = 0; c < 8; ++c) {
        h += tri_w(&t, c, -1) * theta[tri_cell(p, &t, c)];
    }
    return h;
}

/* hd into the grid gradient and hd * dh/dx into gx */
static void size_back(const tn_adj_problem* p, const double* theta, const double* x, double hd, double* gx,
                      double* gtheta) {
    Tri t;
    tri_at(p, x, &t);
    for (int c = 0; c < 8; ++c) {
        const long cell = tri_cell(p, &t, c);
        gtheta[cell] += hd * tri_w(&t, c, -1);
        for (int a = 0; a < 3; ++a) {
            const double s = ((c >> a) & 1) ? 1.0 : -1.0;
            gx[a] += hd * s * tri_w(&t, c, a) * theta[cell] * t.gs[a];
        }
    }
}

static double sp_of(double u, double beta) {
    return u > 0.0 ? u + log1p(exp(-beta * u)) / beta : log1p(exp(beta * u)) / beta;
}

static double sig_of(double v) {
    return 1.0 / (1.0 + exp(-v));
}

static double cfac(const tn_adj_problem* p, int i, int j) {
    return (p->typ[i] != TN_INTERIOR && p->typ[j] != TN_INTERIOR) ? p->fsurf : p->fscale;
}

/* one state: positions x[3n] and sizes h[n] */
static void sizes(const tn_adj_problem* p, const double* theta, const double* x, double* h) {
    for (long i = 0; i < p->n; ++i) {
        h[i] = size_at(p, theta, x + 3 * i);
    }
}

static void step(const tn_adj_problem* p, const double* theta, const double* x, const double* h, double* xn,
                 double* hn) {
    memcpy(xn, x, sizeof(double) * 3 * p->n);
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < p->n; ++i) {
        if (p->typ[i] != TN_INTERIOR) {
            hn[i] = h[i];
            continue;
        }
        double F[3] = { 0, 0, 0 }, a = 0.0;
        for (int k = 0; k < p->nnb[i]; ++k) {
            const int j = p->nbr[i * TN_K + k];
            double e[3];
            for (int d = 0; d < 3; ++d) {
                e[d] = x[3 * i + d] - x[3 * j + d];
            }
            const double l = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
            if (!(l > 0.0)) {
                continue;
            }
            const double u = cfac(p, (int)i, j) * 0.5 * (h[i] + h[j]) - l;
            const double s = sp_of(u, p->beta);
            for (int d = 0; d < 3; ++d) {
                F[d] += s * e[d] / l;
            }
            a += sig_of(p->beta * u);
        }
        for (int d = 0; d < 3; ++d) {
            xn[3 * i + d] = x[3 * i + d] + p->dt * F[d] / (1.0 + a);
        }
        hn[i] = size_at(p, theta, xn + 3 * i);
    }
}

static double tet_q(const tn_adj_problem* p, const double* x, long t, double* V_out, double* Vs_out, double* S_out) {
    const int* v = p->tet + 4 * t;
    double a[3], b[3], c[3];
    for (int d = 0; d < 3; ++d) {
        a[d] = x[3 * v[1] + d] - x[3 * v[0] + d];
        b[d] = x[3 * v[2] + d] - x[3 * v[0] + d];
        c[d] = x[3 * v[3] + d] - x[3 * v[0] + d];
    }
    const double V = (a[0] * (b[1] * c[2] - b[2] * c[1]) + a[1] * (b[2] * c[0] - b[0] * c[2]) +
                      a[2] * (b[0] * c[1] - b[1] * c[0])) / 6.0;
    const double Vs = sqrt(V * V + (double)p->eps_v * p->eps_v);
    double S = 0.0;
    for (int m = 0; m < 4; ++m) {
        for (int o = m + 1; o < 4; ++o) {
            for (int d = 0; d < 3; ++d) {
                const double w = x[3 * v[m] + d] - x[3 * v[o] + d];
                S += w * w;
            }
        }
    }
    *V_out = V;
    *Vs_out = Vs;
    *S_out = S;
    return 12.0 * pow(3.0 * Vs, 2.0 / 3.0) / S;
}

static void loss_at(const tn_adj_problem* p, const double* x, const double* h, double* terms) {
    double E = 0.0, Q = 0.0;
    for (long i = 0; i < p->n; ++i) {
        for (int k = 0; k < p->nnb[i]; ++k) {
            const int j = p->nbr[i * TN_K + k];
            double l2 = 0.0;
            for (int d = 0; d < 3; ++d) {
                const double e = x[3 * i + d] - x[3 * j + d];
                l2 += e * e;
            }
            const double l = sqrt(l2);
            if (!(l > 0.0)) {
                continue;
            }
            const double s = sp_of(cfac(p, (int)i, j) * 0.5 * (h[i] + h[j]) - l, p->beta);
            E += 0.5 * s * s;
        }
    }
    for (long t = 0; t < p->ntet; ++t) {
        double V, Vs, S;
        Q += 1.0 - tet_q(p, x, t, &V, &Vs, &S);
    }
    E = p->n ? E / p->n : 0.0;
    Q = p->ntet ? Q / p->ntet : 0.0;
    terms[0] = E + p->wq * Q;
    terms[1] = E;
    terms[2] = Q;
}

double orc_adj_loss(const tn_adj_problem* p, const double* x0, const double* theta, double* xT, double* terms) {
    const long n = p->n;
    double* x = malloc(sizeof(double) * 3 * n);
    double* y = malloc(sizeof(double) * 3 * n);
    double* h = malloc(sizeof(double) * n);
    double* hy = malloc(sizeof(double) * n);
    memcpy(x, x0, sizeof(double) * 3 * n);
    sizes(p, theta, x, h);
    for (int k = 0; k < p->iters; ++k) {
        step(p, theta, x, h, y, hy);
        double* t = x;
        x = y;
        y = t;
        t = h;
        h = hy;
        hy = t;
    }
    double tm[3];
    loss_at(p, x, h, tm);
    if (xT) {
        memcpy(xT, x, sizeof(double) * 3 * n);
    }
    if (terms) {
        memcpy(terms, tm, sizeof tm);
    }
    free(x);
    free(y);
    free(h);
    free(hy);
    return tm[0];
}

/* d(bar energy or step term)/d(e, l0) pushed to both ends and their sizes */
static void push_bar(long i, long j, const double* ge, double gl0, double c, double* gx, double* hd) {
    for (int d = 0; d < 3; ++d) {
        gx[3 * i + d] += ge[d];
        gx[3 * j + d] -= ge[d];
    }
    hd[i] += 0.5 * c * gl0;
    hd[j] += 0.5 * c * gl0;
}

double orc_adj_grad(const tn_adj_problem* p, const double* x0, const double* theta, double* gx0, double* gtheta,
                    double* terms) {
    const long n = p->n, T = p->iters;
    const long G = (long)p->gx * p->gy * p->gz;
    double* X = malloc(sizeof(double) * 3 * n * (T + 1));   /* every state */
    double* H = malloc(sizeof(double) * n * (T + 1));
    memcpy(X, x0, sizeof(double) * 3 * n);
    sizes(p, theta, X, H);
    for (long k = 0; k < T; ++k) {
        step(p, theta, X + 3 * n * k, H + n * k, X + 3 * n * (k + 1), H + n * (k + 1));
    }
    const double* x = X + 3 * n * T;
    const double* h = H + n * T;
    double tm[3];
    loss_at(p, x, h, tm);
    memset(gtheta, 0, sizeof(double) * G);
    double* lam = calloc(3 * n, sizeof(double));
    double* hd = calloc(n, sizeof(double));
    /* the loss at x_T */
    for (long i = 0; i < n; ++i) {
        for (int k = 0; k < p->nnb[i]; ++k) {
            const int j = p->nbr[i * TN_K + k];
            double e[3];
            for (int d = 0; d < 3; ++d) {
                e[d] = x[3 * i + d] - x[3 * j + d];
            }
            const double l = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
            if (!(l > 0.0)) {
                continue;
            }
            const double c = cfac(p, (int)i, j), u = c * 0.5 * (h[i] + h[j]) - l;
            const double w = sp_of(u, p->beta) * sig_of(p->beta * u) / n;   /* dE/du */
            const double ge[3] = { -w * e[0] / l, -w * e[1] / l, -w * e[2] / l };
            push_bar(i, j, ge, w, c, lam, hd);
        }
    }
    for (long t = 0; t < p->ntet; ++t) {
        const int* v = p->tet + 4 * t;
        double V, Vs, S;
        const double q = tet_q(p, x, t, &V, &Vs, &S);
        const double wt = -p->wq / (double)p->ntet;
        double a[3], b[3], c[3];
        for (int d = 0; d < 3; ++d) {
            a[d] = x[3 * v[1] + d] - x[3 * v[0] + d];
            b[d] = x[3 * v[2] + d] - x[3 * v[0] + d];
            c[d] = x[3 * v[3] + d] - x[3 * v[0] + d];
        }
        double dv[4][3];
        dv[1][0] = b[1] * c[2] - b[2] * c[1];
        dv[1][1] = b[2] * c[0] - b[0] * c[2];
        dv[1][2] = b[0] * c[1] - b[1] * c[0];
        dv[2][0] = c[1] * a[2] - c[2] * a[1];
        dv[2][1] = c[2] * a[0] - c[0] * a[2];
        dv[2][2] = c[0] * a[1] - c[1] * a[0];
        dv[3][0] = a[1] * b[2] - a[2] * b[1];
        dv[3][1] = a[2] * b[0] - a[0] * b[2];
        dv[3][2] = a[0] * b[1] - a[1] * b[0];
        for (int d = 0; d < 3; ++d) {
            dv[0][d] = -(dv[1][d] + dv[2][d] + dv[3][d]);
        }
        const double kv = q * (2.0 / 3.0) * (V / (Vs * Vs)) / 6.0, ks = q / S;
        for (int m = 0; m < 4; ++m) {
            for (int d = 0; d < 3; ++d) {
                double ds = 0.0;
                for (int o = 0; o < 4; ++o) {
                    if (o != m) {
                        ds += 2.0 * (x[3 * v[m] + d] - x[3 * v[o] + d]);
                    }
                }
                lam[3 * v[m] + d] += wt * (kv * dv[m][d] - ks * ds);
            }
        }
    }
    for (long i = 0; i < n; ++i) {
        size_back(p, theta, x + 3 * i, hd[i], lam + 3 * i, gtheta);
    }
    /* the steps, backwards: x' = x + m dt F / D */
    double* nl = malloc(sizeof(double) * 3 * n);
    for (long k = T - 1; k >= 0; --k) {
        const double* xk = X + 3 * n * k;
        const double* hk = H + n * k;
        memcpy(nl, lam, sizeof(double) * 3 * n);   /* the identity part */
        memset(hd, 0, sizeof(double) * n);
        for (long i = 0; i < n; ++i) {
            if (p->typ[i] != TN_INTERIOR) {
                continue;
            }
            double F[3] = { 0, 0, 0 }, a = 0.0;
            for (int kk = 0; kk < p->nnb[i]; ++kk) {
                const int j = p->nbr[i * TN_K + kk];
                double e[3];
                for (int d = 0; d < 3; ++d) {
                    e[d] = xk[3 * i + d] - xk[3 * j + d];
                }
                const double l = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
                if (!(l > 0.0)) {
                    continue;
                }
                const double u = cfac(p, (int)i, j) * 0.5 * (hk[i] + hk[j]) - l;
                const double s = sp_of(u, p->beta);
                for (int d = 0; d < 3; ++d) {
                    F[d] += s * e[d] / l;
                }
                a += sig_of(p->beta * u);
            }
            const double D = 1.0 + a;
            /* x'_i = x_i + dt F / D: d x'_i = dt (dF / D - F dD / D^2) */
            double nu[3], kap = 0.0;
            for (int d = 0; d < 3; ++d) {
                nu[d] = p->dt * lam[3 * i + d] / D;
                kap += p->dt * lam[3 * i + d] * F[d];
            }
            kap /= D * D;
            for (int kk = 0; kk < p->nnb[i]; ++kk) {
                const int j = p->nbr[i * TN_K + kk];
                double e[3];
                for (int d = 0; d < 3; ++d) {
                    e[d] = xk[3 * i + d] - xk[3 * j + d];
                }
                const double l = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
                if (!(l > 0.0)) {
                    continue;
                }
                const double c = cfac(p, (int)i, j), u = c * 0.5 * (hk[i] + hk[j]) - l;
                const double s = sp_of(u, p->beta), sg = sig_of(p->beta * u), rho = p->beta * sg * (1.0 - sg);
                double nh[3];
                for (int d = 0; d < 3; ++d) {
                    nh[d] = e[d] / l;
                }
                const double vn = nu[0] * nh[0] + nu[1] * nh[1] + nu[2] * nh[2];
                /* nu . d(s e / l)/de = (s / l)(nu - vn nh) - sg vn nh; - kap d sg/de = kap rho nh */
                double ge[3];
                for (int d = 0; d < 3; ++d) {
                    ge[d] = (s / l) * (nu[d] - vn * nh[d]) - sg * vn * nh[d] + kap * rho * nh[d];
                }
                push_bar(i, j, ge, sg * vn - kap * rho, c, nl, hd);
            }
        }
        for (long i = 0; i < n; ++i) {
            size_back(p, theta, xk + 3 * i, hd[i], nl + 3 * i, gtheta);
        }
        memcpy(lam, nl, sizeof(double) * 3 * n);
    }
    memcpy(gx0, lam, sizeof(double) * 3 * n);
    if (terms) {
        memcpy(terms, tm, sizeof tm);
    }
    free(X);
    free(H);
    free(lam);
    free(hd);
    free(nl);
    return tm[0];
}
