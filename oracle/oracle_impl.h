/*
 * oracle_impl.h -- precision-generic body of the CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Included twice by oracle.c, once with REAL=float (suffix _f32) and once with
 * REAL=double (suffix _f64).  It restates, step for step and in the same operation
 * order, the reference CDL ADMM of CDLmotif.jl:
 *   src/SETUP/CDL_setup.jl:15-40,165-236,285-286   (state, init, objective)
 *   src/CDL/train.jl:1-98                           (code_update!, dict_update!, update_best!)
 *   src/CDL/csc_kernels.jl:8-95, cdl_kernels.jl:2-77 (device kernels)
 *   src/WRAPPER/wrap.jl:12                          (code threshold)
 * Every transform is a separate FFT of one array, as in train.jl ("reference-structured").
 * The FFT is this file's own mixed-radix decimation-in-time transform (the reference uses
 * cuFFT C2C plans, CDL_setup.jl:158-163,195-204; the inverse is the unscaled inverse times
 * REAL(1/L_gamma), the AbstractFFTs ScaledPlan convention).
 *
 * Nothing in the product links this file; only tests, the smoke check and the CPU leg of
 * the benchmark load it, as the checker.
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SFX)

typedef struct { REAL re, im; } FN(cplx_);
#define CPLX FN(cplx_)

static inline CPLX FN(cmul)(CPLX a, CPLX b) { CPLX r; r.re = a.re * b.re - a.im * b.im; r.im = a.re * b.im + a.im * b.re; return r; }
static inline CPLX FN(cadd)(CPLX a, CPLX b) { CPLX r; r.re = a.re + b.re; r.im = a.im + b.im; return r; }
static inline CPLX FN(csub)(CPLX a, CPLX b) { CPLX r; r.re = a.re - b.re; r.im = a.im - b.im; return r; }
static inline CPLX FN(cconj)(CPLX a) { CPLX r; r.re = a.re; r.im = -a.im; return r; }
/* Julia's Complex / Complex (Base.:/ for Complex{Float32} goes through the Float64 path
 * for Float32 inputs and rounds back); the divisors here have zero imaginary part. */
static inline CPLX FN(cdiv)(CPLX a, CPLX b) {
    double den = (double)b.re * b.re + (double)b.im * b.im;
    CPLX r;
    r.re = (REAL)(((double)a.re * b.re + (double)a.im * b.im) / den);
    r.im = (REAL)(((double)a.im * b.re - (double)a.re * b.im) / den);
    return r;
}

/* ---------------- FFT: mixed radix, any length ---------------- */
typedef struct {
    int n, nf, f[40];
    REAL *wre, *wim;   /* W_n^t = exp(-2*pi*i*t/n), rounded from double */
} FN(fftplan_);
#define FFTPLAN FN(fftplan_)

static void FN(fft_plan_init)(FFTPLAN *p, int n) {
    p->n = n; p->nf = 0;
    int m = n;
    for (int q = 2; m > 1;) {
        if (m % q == 0) { p->f[p->nf++] = q; m /= q; } else q++;
    }
    p->wre = (REAL *)malloc(sizeof(REAL) * n);
    p->wim = (REAL *)malloc(sizeof(REAL) * n);
    for (int t = 0; t < n; t++) {
        double a = -2.0 * M_PI * (double)t / (double)n;
        p->wre[t] = (REAL)cos(a);
        p->wim[t] = (REAL)sin(a);
    }
}
static void FN(fft_plan_free)(FFTPLAN *p) { free(p->wre); free(p->wim); }

static inline CPLX FN(tw)(const FFTPLAN *P, long t, int inv) {
    t %= P->n; if (t < 0) t += P->n;
    CPLX w; w.re = P->wre[t]; w.im = inv ? -P->wim[t] : P->wim[t];
    return w;
}

/* out[0..n) = DFT of in[0], in[istride], ... ; fi indexes the factor list. */
static void FN(fft_rec)(const FFTPLAN *P, int n, int fi, const CPLX *in, long istride, CPLX *out, int inv) {
    if (n == 1) { out[0] = in[0]; return; }
    int p = P->f[fi], m = n / p;
    long s = P->n / n;                        /* W_n^t = W_N^(t*s) */
    for (int q = 0; q < p; q++) FN(fft_rec)(P, m, fi + 1, in + q * istride, istride * p, out + (long)q * m, inv);
    CPLX t[128], y;
    for (int k = 0; k < m; k++) {
        for (int q = 0; q < p; q++) t[q] = FN(cmul)(out[(long)q * m + k], FN(tw)(P, (long)q * k * s, inv));
        for (int r = 0; r < p; r++) {
            y = t[0];
            for (int q = 1; q < p; q++) y = FN(cadd)(y, FN(cmul)(t[q], FN(tw)(P, (long)((q * r) % p) * (P->n / p), inv)));
            out[k + (long)r * m] = y;
        }
    }
}

/* forward, unscaled */
static void FN(fft_c)(const FFTPLAN *P, const CPLX *in, CPLX *out) { FN(fft_rec)(P, P->n, 0, in, 1, out, 0); }
/* inverse with the ScaledPlan convention: unscaled inverse, then times REAL(1/n) */
static void FN(ifft_c)(const FFTPLAN *P, const CPLX *in, CPLX *out) {
    FN(fft_rec)(P, P->n, 0, in, 1, out, 1);
    REAL sc = (REAL)(1.0 / (double)P->n);
    for (int i = 0; i < P->n; i++) { out[i].re *= sc; out[i].im *= sc; }
}
static void FN(fft_r)(const FFTPLAN *P, const REAL *x, CPLX *out, CPLX *tmp) {
    for (int i = 0; i < P->n; i++) { tmp[i].re = x[i]; tmp[i].im = 0; }
    FN(fft_c)(P, tmp, out);
}

/* Batched helpers over contiguous columns of length n. */
static void FN(fft_cols)(const FFTPLAN *P, const REAL *x, CPLX *xh, long ncol) {
    int n = P->n;
#pragma omp parallel
    {
        CPLX *tmp = (CPLX *)malloc(sizeof(CPLX) * n);
#pragma omp for schedule(static)
        for (long c = 0; c < ncol; c++) FN(fft_r)(P, x + c * n, xh + c * n, tmp);
        free(tmp);
    }
}
static void FN(ifft_cols_real)(const FFTPLAN *P, const CPLX *xh, REAL *x, long ncol) {
    int n = P->n;
#pragma omp parallel
    {
        CPLX *tmp = (CPLX *)malloc(sizeof(CPLX) * n);
#pragma omp for schedule(static)
        for (long c = 0; c < ncol; c++) {
            FN(ifft_c)(P, xh + c * n, tmp);
            for (int i = 0; i < n; i++) x[c * n + i] = tmp[i].re;
        }
        free(tmp);
    }
}

/* ---------------- CDL state (CDL_setup.jl:70-164) ---------------- */
typedef struct {
    int L4, N, M, Lg, gamma;
    int *fl;
    REAL rho, sigma, lambda;
    REAL *s;                       /* (L4, N) data */
    REAL *P;                       /* (Lg, M) projector, CDL_setup.jl:15-27 */
    REAL *WtW;                     /* diag(W'W), (Lg) */
    REAL *WtS;                     /* (Lg, N) */
    REAL *X, *d, *Td, *DX;
    CPLX *Xh, *dh;
    REAL obj_best;
    REAL *Dbest, *Xbest;
    REAL *Y1, *Y2, *Y3, *y1, *y2, *y3;
    CPLX *Y1h, *Y2h, *Y3h, *y1h, *y2h, *y3h;
    REAL *G1, *G2, *G3, *g1, *g2, *g3;
    CPLX *G1h, *G2h, *G3h, *g1h, *g2h, *g3h;
    FFTPLAN plan;
} FN(cdlstate_);
#define STATE FN(cdlstate_)

#define NEWR(cnt) ((REAL *)calloc((size_t)(cnt), sizeof(REAL)))
#define NEWC(cnt) ((CPLX *)calloc((size_t)(cnt), sizeof(CPLX)))

/* 0.5*norm(W*DX .- s)^2 + lambda*reduce(+, abs.(X))   (CDL_setup.jl:286)
 * Julia types: norm -> REAL, ^2 in REAL, 0.5 (Float64) * that -> Float64; lambda*sum in REAL.
 * The sums are taken in double and rounded once (the reference's GPU reduction order is
 * not specified). */
static double FN(objective_of)(const STATE *c, const REAL *X, const REAL *DX) {
    double acc = 0.0;
    for (long n = 0; n < c->N; n++)
        for (int i = 0; i < c->L4; i++) {
            double r = (double)(REAL)(DX[n * c->Lg + c->gamma + i] - c->s[n * c->L4 + i]);
            acc += r * r;
        }
    REAL nrm = (REAL)sqrt(acc);
    REAL nrm2 = nrm * nrm;
    double ab = 0.0;
    long E = (long)c->Lg * c->M * c->N;
    for (long e = 0; e < E; e++) ab += fabs((double)X[e]);
    REAL reg = c->lambda * (REAL)ab;
    return 0.5 * (double)nrm2 + (double)reg;
}

/* projection2box! (cdl_kernels.jl:37-51): out = P!=0 ? clamp(v,0,1) : 0, per (l, m). */
void FN(orc_box)(int Lg, int M, const REAL *v, const REAL *P, REAL *out) {
    for (long i = 0; i < (long)Lg * M; i++) {
        REAL x = v[i];
        if (x >= 1) out[i] = (P[i] != 0) ? (REAL)1 : (REAL)0;
        else if (x <= 0) out[i] = 0;
        else out[i] = (P[i] != 0) ? x : (REAL)0;
    }
}

/* prox_simplex_zeros! (cdl_kernels.jl:53-66): on the support,
 * out[l] = v[l] - 0.25*(v[i]+v[i+1]+v[i+2]+v[i+3] - 1) with i = 4*floor(l_1based/4)
 * (0-based start; quirk: rows with l_1based % 4 == 0 use the next group); the group sum
 * is formed in REAL, the combination in double (Float64 literal 0.25), stored as REAL. */
void FN(orc_prox_simplex)(int Lg, int M, const REAL *vavg, const REAL *P, REAL *out) {
    for (int m = 0; m < M; m++)
        for (int l0 = 0; l0 < Lg; l0++) {
            long i = (long)m * Lg + l0;
            if (P[i] != 0) {
                int gi = 4 * ((l0 + 1) / 4);
                const REAL *v = vavg + (long)m * Lg;
                REAL s4 = v[gi] + v[gi + 1] + v[gi + 2] + v[gi + 3] - (REAL)1;
                out[i] = (REAL)((double)v[l0] - 0.25 * (double)s4);
            } else out[i] = 0;
        }
}

STATE *FN(orc_cdl_create)(int L4, int N, int M, const int *fl, double rho, double sigma, double lambda,
                          const float *s, const float *D0, const float *X0) {
    STATE *c = (STATE *)calloc(1, sizeof(STATE));
    c->L4 = L4; c->N = N; c->M = M;
    c->fl = (int *)malloc(sizeof(int) * M);
    int gamma = 0;
    for (int m = 0; m < M; m++) { c->fl[m] = fl[m]; if (fl[m] > gamma) gamma = fl[m]; }
    c->gamma = gamma;
    int Lg = L4 + 2 * gamma;                     /* CDL_setup.jl:180 */
    c->Lg = Lg;
    c->rho = (REAL)rho; c->sigma = (REAL)sigma; c->lambda = (REAL)lambda;
    long LM = (long)Lg * M, LN = (long)Lg * N, E = LM * N;
    FN(fft_plan_init)(&c->plan, Lg);

    c->s = NEWR((long)L4 * N);
    for (long i = 0; i < (long)L4 * N; i++) c->s[i] = (REAL)s[i];
    c->P = NEWR(LM);
    for (int m = 0; m < M; m++) for (int l = 0; l < Lg; l++) c->P[(long)m * Lg + l] = (l < fl[m]) ? 1 : 0;
    REAL *D = NEWR(LM);
    for (long i = 0; i < LM; i++) D[i] = c->P[i] * (REAL)D0[i];          /* :183 */

    c->X = NEWR(E);
    for (long e = 0; e < E; e++) {
        int l = (int)(e % Lg);
        c->X[e] = (l % 4 != 0) ? (REAL)0 : (REAL)X0[e];              /* :191-192 */
    }
    CPLX *Dh = NEWC(LM);
    FN(fft_cols)(&c->plan, D, Dh, M);                                  /* :197 */
    c->Xh = NEWC(E);
    FN(fft_cols)(&c->plan, c->X, c->Xh, (long)M * N);                  /* :198 */
    CPLX *DoX = NEWC(LN);
    for (long n = 0; n < N; n++)                                        /* :201 */
        for (int l = 0; l < Lg; l++) {
            CPLX acc = {0, 0};
            for (int m = 0; m < M; m++) acc = FN(cadd)(acc, FN(cmul)(Dh[(long)m * Lg + l], c->Xh[((long)n * M + m) * Lg + l]));
            DoX[n * Lg + l] = acc;
        }
    c->DX = NEWR(LN);
    FN(ifft_cols_real)(&c->plan, DoX, c->DX, N);                        /* :203 */
    free(DoX);

    c->dh = NEWC(E);                                                     /* :207-208 */
    for (long n = 0; n < N; n++) memcpy(c->dh + (long)n * LM, Dh, sizeof(CPLX) * LM);
    c->Td = NEWR(LN);                                                    /* :209 */
    c->d = NEWR(E);
    FN(ifft_cols_real)(&c->plan, c->dh, c->d, (long)M * N);             /* :210 */

    c->WtW = NEWR(Lg);                                                   /* :213-217 */
    for (int l = 0; l < Lg; l++) c->WtW[l] = (l >= gamma && l < gamma + L4) ? 1 : 0;
    c->WtS = NEWR(LN);
    for (long n = 0; n < N; n++) for (int i = 0; i < L4; i++) c->WtS[n * Lg + gamma + i] = c->s[n * L4 + i];

    c->Y1 = NEWR(LN); c->Y2 = NEWR(E); c->Y3 = NEWR(E);                 /* :220 */
    memcpy(c->Y2, c->X, sizeof(REAL) * E); memcpy(c->Y3, c->X, sizeof(REAL) * E);
    c->Y1h = NEWC(LN); c->Y2h = NEWC(E); c->Y3h = NEWC(E);               /* :221 */
    FN(fft_cols)(&c->plan, c->Y2, c->Y2h, (long)M * N);
    memcpy(c->Y3h, c->Y2h, sizeof(CPLX) * E);
    c->y1 = NEWR(LN); c->y2 = NEWR(E); c->y3 = NEWR(E);                 /* :223-224 */
    c->y1h = NEWC(LN); c->y2h = NEWC(E); c->y3h = NEWC(E);

    c->G1 = NEWR(LM); c->G2 = NEWR(LN); c->G3 = NEWR(LM);              /* :226-227 */
    memcpy(c->G1, D, sizeof(REAL) * LM); memcpy(c->G2, c->DX, sizeof(REAL) * LN); memcpy(c->G3, D, sizeof(REAL) * LM);
    c->G1h = NEWC(LM); c->G2h = NEWC(LN); c->G3h = NEWC(LM);
    FN(fft_cols)(&c->plan, c->G1, c->G1h, M);
    FN(fft_cols)(&c->plan, c->G2, c->G2h, N);
    memcpy(c->G3h, c->G1h, sizeof(CPLX) * LM);
    c->g1 = NEWR(E); c->g2 = NEWR(LN); c->g3 = NEWR(E);                 /* :229-230 */
    c->g1h = NEWC(E); c->g2h = NEWC(LN); c->g3h = NEWC(E);

    c->Dbest = NEWR(LM); memcpy(c->Dbest, c->G1, sizeof(REAL) * LM);   /* :233-234 */
    c->Xbest = NEWR(E); memcpy(c->Xbest, c->Y3, sizeof(REAL) * E);
    c->obj_best = (REAL)FN(objective_of)(c, c->X, c->DX);              /* :236, stored as REAL */
    free(D); free(Dh);
    return c;
}

void FN(orc_cdl_destroy)(STATE *c) {
    if (!c) return;
    void *ptrs[] = {c->fl, c->s, c->P, c->WtW, c->WtS, c->X, c->d, c->Td, c->DX, c->Xh, c->dh, c->Dbest, c->Xbest,
                    c->Y1, c->Y2, c->Y3, c->y1, c->y2, c->y3, c->Y1h, c->Y2h, c->Y3h, c->y1h, c->y2h, c->y3h,
                    c->G1, c->G2, c->G3, c->g1, c->g2, c->g3, c->G1h, c->G2h, c->G3h, c->g1h, c->g2h, c->g3h};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
    FN(fft_plan_free)(&c->plan);
    free(c);
}

/* train.jl:1-44 */
void FN(orc_cdl_code_update)(STATE *c, int num_iter, double sherman) {
    const int Lg = c->Lg, M = c->M, N = c->N;
    const long LM = (long)Lg * M, LN = (long)Lg * N, E = LM * N;
    const REAL shc = (REAL)sherman;
    CPLX *Ah = NEWC(LM), *DXh = NEWC(LN);
    REAL *sumd = NEWR(Lg);
    CPLX *alpha = NEWC(LN);
    for (int it = 0; it < num_iter; it++) {
        /* :7 shrink! */
        for (long e = 0; e < E; e++) {
            REAL t = c->X[e] + c->y2[e];
            REAL a = fabs(t) - c->lambda;
            REAL sg = (t > 0) ? (REAL)1 : ((t < 0) ? (REAL)-1 : (REAL)0);
            c->Y2[e] = sg * (a > 0 ? a : (REAL)0);
        }
        FN(fft_cols)(&c->plan, c->Y2, c->Y2h, (long)M * N);               /* :8 */
        for (long i = 0; i < LN; i++) c->y1[i] = c->y1[i] + c->DX[i] - c->Y1[i];   /* :12 */
        FN(fft_cols)(&c->plan, c->y1, c->y1h, N);
        for (long e = 0; e < E; e++) c->y2[e] = c->y2[e] + c->X[e] - c->Y2[e];     /* :13 */
        FN(fft_cols)(&c->plan, c->y2, c->y2h, (long)M * N);
        for (long e = 0; e < E; e++) c->y3[e] = c->y3[e] + c->X[e] - c->Y3[e];     /* :14 */
        FN(fft_cols)(&c->plan, c->y3, c->y3h, (long)M * N);
        /* :18-20  dots = G1h .* conj(G1h); sum over m (+ sherman); A = G1h ./ sum */
        for (int l = 0; l < Lg; l++) {
            REAL acc = 0;
            for (int m = 0; m < M; m++) {
                CPLX g = c->G1h[(long)m * Lg + l];
                CPLX p = FN(cmul)(g, FN(cconj)(g));
                acc += p.re;
            }
            sumd[l] = acc + shc;
        }
        for (int m = 0; m < M; m++)
            for (int l = 0; l < Lg; l++) {
                CPLX den = {sumd[l], 0};
                Ah[(long)m * Lg + l] = FN(cdiv)(c->G1h[(long)m * Lg + l], den);
            }
        /* :23-25 RHS, dot, combine (per n; RHS kept per column) */
#pragma omp parallel
        {
            CPLX *rhs = NEWC(LM);
#pragma omp for schedule(static)
            for (long n = 0; n < N; n++) {
                for (int m = 0; m < M; m++)
                    for (int l = 0; l < Lg; l++) {
                        long e = ((long)n * M + m) * Lg + l;
                        CPLX g = c->G1h[(long)m * Lg + l];
                        CPLX a = FN(cmul)(FN(cconj)(g), FN(csub)(c->Y1h[n * Lg + l], c->y1h[n * Lg + l]));
                        a = FN(cadd)(a, FN(csub)(c->Y2h[e], c->y2h[e]));
                        a = FN(cadd)(a, FN(csub)(c->Y3h[e], c->y3h[e]));
                        rhs[(long)m * Lg + l] = a;                           /* csc_kernels.jl:12 */
                    }
                for (int l = 0; l < Lg; l++) {                                /* csc_kernels.jl:27-30 */
                    CPLX acc = {0, 0};
                    for (int m = 0; m < M; m++) acc = FN(cadd)(acc, FN(cmul)(Ah[(long)m * Lg + l], rhs[(long)m * Lg + l]));
                    alpha[n * Lg + l] = acc;
                }
                for (int m = 0; m < M; m++)                                   /* csc_kernels.jl:44 */
                    for (int l = 0; l < Lg; l++) {
                        long e = ((long)n * M + m) * Lg + l;
                        CPLX t = FN(csub)(rhs[(long)m * Lg + l], FN(cmul)(alpha[n * Lg + l], FN(cconj)(c->G1h[(long)m * Lg + l])));
                        t.re = t.re / shc; t.im = t.im / shc;
                        c->Xh[e] = t;
                    }
            }
            free(rhs);
        }
        FN(ifft_cols_real)(&c->plan, c->Xh, c->X, (long)M * N);          /* :26 */
        for (long n = 0; n < N; n++)                                      /* :27 */
            for (int l = 0; l < Lg; l++) {
                CPLX acc = {0, 0};
                for (int m = 0; m < M; m++) acc = FN(cadd)(acc, FN(cmul)(c->G1h[(long)m * Lg + l], c->Xh[((long)n * M + m) * Lg + l]));
                DXh[n * Lg + l] = acc;
            }
        FN(ifft_cols_real)(&c->plan, DXh, c->DX, N);
        for (long n = 0; n < N; n++)                                      /* :35 update_Y1! */
            for (int l = 0; l < Lg; l++) {
                long i = n * Lg + l;
                c->Y1[i] = (c->WtS[i] + c->rho * (c->DX[i] + c->y1[i])) / (c->WtW[l] + c->rho);
            }
        FN(fft_cols)(&c->plan, c->Y1, c->Y1h, N);                          /* :36 */
        for (long e = 0; e < E; e++) {                                     /* :40 projectionX! */
            int l = (int)(e % Lg);
            c->Y3[e] = (l % 4 != 0) ? (REAL)0 : c->X[e] + c->y3[e];
        }
        FN(fft_cols)(&c->plan, c->Y3, c->Y3h, (long)M * N);               /* :41 */
    }
    free(Ah); free(DXh); free(sumd); free(alpha);
}

/* train.jl:46-89 */
void FN(orc_cdl_dict_update)(STATE *c, int num_iter, double sherman) {
    const int Lg = c->Lg, M = c->M, N = c->N;
    const long LM = (long)Lg * M, LN = (long)Lg * N, E = LM * N;
    const REAL shc = (REAL)sherman;
    CPLX *Tdh = NEWC(LN);
    REAL *vavg = NEWR(LM);
    for (int it = 0; it < num_iter; it++) {
#pragma omp parallel
        {
            CPLX *Xc = NEWC(LM), *rhs = NEWC(LM);
            REAL *sd = NEWR(Lg);
#pragma omp for schedule(static)
            for (long n = 0; n < N; n++) {
                const CPLX *Y3h = c->Y3h + (long)n * LM;
                for (int l = 0; l < Lg; l++) {                                 /* :52-53 */
                    REAL acc = 0;
                    for (int m = 0; m < M; m++) { CPLX y = Y3h[(long)m * Lg + l]; acc += FN(cmul)(y, FN(cconj)(y)).re; }
                    sd[l] = acc + shc;
                }
                for (int m = 0; m < M; m++)                                    /* :54 */
                    for (int l = 0; l < Lg; l++) {
                        CPLX den = {sd[l], 0};
                        Xc[(long)m * Lg + l] = FN(cdiv)(Y3h[(long)m * Lg + l], den);
                    }
                for (int l = 0; l < Lg; l++) {                                 /* :56 (old dh) */
                    CPLX acc = {0, 0};
                    for (int m = 0; m < M; m++) acc = FN(cadd)(acc, FN(cmul)(Y3h[(long)m * Lg + l], c->dh[(long)n * LM + (long)m * Lg + l]));
                    Tdh[n * Lg + l] = acc;
                }
                for (int m = 0; m < M; m++)                                    /* :58 sherman_CDL_RHS! */
                    for (int l = 0; l < Lg; l++) {
                        long e = (long)n * LM + (long)m * Lg + l;
                        CPLX a = FN(cmul)(FN(cconj)(Y3h[(long)m * Lg + l]), FN(csub)(c->G2h[n * Lg + l], c->g2h[n * Lg + l]));
                        a = FN(cadd)(a, FN(csub)(c->G1h[(long)m * Lg + l], c->g1h[e]));
                        a = FN(cadd)(a, FN(csub)(c->G3h[(long)m * Lg + l], c->g3h[e]));
                        rhs[(long)m * Lg + l] = a;
                    }
                for (int l = 0; l < Lg; l++) {                                 /* :59 beta, :60 combine (no 1/2) */
                    CPLX beta = {0, 0};
                    for (int m = 0; m < M; m++) beta = FN(cadd)(beta, FN(cmul)(Xc[(long)m * Lg + l], rhs[(long)m * Lg + l]));
                    for (int m = 0; m < M; m++) {
                        long e = (long)n * LM + (long)m * Lg + l;
                        c->dh[e] = FN(csub)(rhs[(long)m * Lg + l], FN(cmul)(beta, FN(cconj)(Y3h[(long)m * Lg + l])));
                    }
                }
            }
            free(Xc); free(rhs); free(sd);
        }
        FN(ifft_cols_real)(&c->plan, Tdh, c->Td, N);                        /* :56-57 */
        FN(ifft_cols_real)(&c->plan, c->dh, c->d, (long)M * N);             /* :61 */
        for (long n = 0; n < N; n++)                                         /* :68 update_g2! */
            for (int l = 0; l < Lg; l++) {
                long i = n * Lg + l;
                c->G2[i] = (c->WtS[i] + c->sigma * (c->Td[i] + c->g2[i])) / (c->WtW[l] + c->sigma);
            }
        FN(fft_cols)(&c->plan, c->G2, c->G2h, N);                             /* :69 */
        for (long e = 0; e < E; e++) c->g1[e] = c->g1[e] + c->d[e] - c->G1[e % LM];   /* :70 */
        FN(fft_cols)(&c->plan, c->g1, c->g1h, (long)M * N);
        for (long i = 0; i < LN; i++) c->g2[i] = c->g2[i] + c->Td[i] - c->G2[i];     /* :71 */
        FN(fft_cols)(&c->plan, c->g2, c->g2h, N);
        for (long e = 0; e < E; e++) c->g3[e] = c->g3[e] + c->d[e] - c->G3[e % LM];   /* :72 */
        FN(fft_cols)(&c->plan, c->g3, c->g3h, (long)M * N);
        /* :76-78 v_avg_2 = sum_n(d + g3)/N; projection2box! */
        for (long i = 0; i < LM; i++) {
            REAL acc = 0;
            for (long n = 0; n < N; n++) acc += c->d[n * LM + i] + c->g3[n * LM + i];
            vavg[i] = acc / (REAL)N;
        }
        FN(orc_box)(Lg, M, vavg, c->P, c->G3);
        FN(fft_cols)(&c->plan, c->G3, c->G3h, M);
        /* :83-85 v_avg_1 = sum_n(d + g1)/N; prox_simplex_zeros! */
        for (long i = 0; i < LM; i++) {
            REAL acc = 0;
            for (long n = 0; n < N; n++) acc += c->d[n * LM + i] + c->g1[n * LM + i];
            vavg[i] = acc / (REAL)N;
        }
        FN(orc_prox_simplex)(Lg, M, vavg, c->P, c->G1);
        FN(fft_cols)(&c->plan, c->G1, c->G1h, M);
    }
    free(Tdh); free(vavg);
}

double FN(orc_cdl_objective)(STATE *c) { return FN(objective_of)(c, c->X, c->DX); }

/* train.jl:91-98.  Returns 1 when the best iterate was replaced. */
int FN(orc_cdl_update_best)(STATE *c, double *obj_out) {
    double obj = FN(objective_of)(c, c->X, c->DX);
    if (obj_out) *obj_out = obj;
    if (obj < (double)c->obj_best) {
        const long LM = (long)c->Lg * c->M, E = LM * c->N;
        c->obj_best = (REAL)obj;
        for (long i = 0; i < LM; i++) {
            REAL acc = 0;
            for (long n = 0; n < c->N; n++) acc += c->d[n * LM + i];
            c->Dbest[i] = acc / (REAL)c->N;
        }
        memcpy(c->Xbest, c->X, sizeof(REAL) * E);
        return 1;
    }
    return 0;
}

double FN(orc_cdl_obj_best)(STATE *c) { return (double)c->obj_best; }

/* wrap.jl:12 -- mult * std(X_best[X_best .> 0]) (corrected), statistics in double */
double FN(orc_cdl_code_thresh)(STATE *c, double mult, long *count_out) {
    const long E = (long)c->Lg * c->M * c->N;
    long cnt = 0; double s1 = 0;
    for (long e = 0; e < E; e++) if (c->Xbest[e] > 0) { cnt++; s1 += (double)c->Xbest[e]; }
    double mean = cnt ? s1 / (double)cnt : 0.0, ss = 0;
    for (long e = 0; e < E; e++) if (c->Xbest[e] > 0) { double t = (double)c->Xbest[e] - mean; ss += t * t; }
    if (count_out) *count_out = cnt;
    if (cnt < 2) return NAN;
    return mult * (double)(REAL)sqrt(ss / (double)(cnt - 1));
}

/* Field copy-out. */
#ifndef ORACLE_FIELDS_DEFINED
#define ORACLE_FIELDS_DEFINED
enum { OF_X = 0, OF_DX, OF_D, OF_G1, OF_G2, OF_G3, OF_XBEST, OF_DBEST, OF_Y1, OF_Y2, OF_Y3,
       OF_y1, OF_y2, OF_y3, OF_g1, OF_g2, OF_g3, OF_TD };
#endif
long FN(orc_cdl_get)(STATE *c, int field, REAL *dst) {
    const long LM = (long)c->Lg * c->M, LN = (long)c->Lg * c->N, E = LM * c->N;
    const REAL *src = NULL; long cnt = 0;
    switch (field) {
    case OF_X: src = c->X; cnt = E; break;
    case OF_DX: src = c->DX; cnt = LN; break;
    case OF_D: src = c->d; cnt = E; break;
    case OF_G1: src = c->G1; cnt = LM; break;
    case OF_G2: src = c->G2; cnt = LN; break;
    case OF_G3: src = c->G3; cnt = LM; break;
    case OF_XBEST: src = c->Xbest; cnt = E; break;
    case OF_DBEST: src = c->Dbest; cnt = LM; break;
    case OF_Y1: src = c->Y1; cnt = LN; break;
    case OF_Y2: src = c->Y2; cnt = E; break;
    case OF_Y3: src = c->Y3; cnt = E; break;
    case OF_y1: src = c->y1; cnt = LN; break;
    case OF_y2: src = c->y2; cnt = E; break;
    case OF_y3: src = c->y3; cnt = E; break;
    case OF_g1: src = c->g1; cnt = E; break;
    case OF_g2: src = c->g2; cnt = LN; break;
    case OF_g3: src = c->g3; cnt = E; break;
    case OF_TD: src = c->Td; cnt = LN; break;
    default: return -1;
    }
    if (dst) memcpy(dst, src, sizeof(REAL) * cnt);
    return cnt;
}

/* Exposed for the FFT checks: forward (inv=0) or scaled inverse (inv=1) complex transform. */
void FN(orc_fft)(int n, const REAL *re_in, const REAL *im_in, REAL *re_out, REAL *im_out, int inv) {
    FFTPLAN P; FN(fft_plan_init)(&P, n);
    CPLX *a = NEWC(n), *b = NEWC(n);
    for (int i = 0; i < n; i++) { a[i].re = re_in[i]; a[i].im = im_in[i]; }
    if (inv) FN(ifft_c)(&P, a, b); else FN(fft_c)(&P, a, b);
    for (int i = 0; i < n; i++) { re_out[i] = b[i].re; im_out[i] = b[i].im; }
    free(a); free(b); FN(fft_plan_free)(&P);
}

#undef CPLX
#undef FFTPLAN
#undef STATE
#undef NEWR
#undef NEWC
