/*
 * CPU oracle for the iso-acoustic forward / adjoint hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker the HIP library is
 * compared against (and the CPU baseline timed beside it). It is never linked
 * into, called by, or used as a fallback for the product library.
 *
 * It restates, in plain C over padded float arrays, the discrete equations
 * that Stride hands to Devito for the second-order iso-acoustic wave equation
 * (stride/physics/iso_acoustic/devito.py) and the boundary conditions
 * (stride/physics/boundaries/devito.py):
 *
 *   - SO10 centred Laplacian, or vp^2 * div(grad) composed of SO10 staggered
 *     first derivatives for constant / variable density        devito.py:1646-1666
 *   - OT2 single pass, OT4 two-pass (L2 then L(L2))             devito.py:1490-1534, 1632-1644
 *   - attenuation power 0 (centred p_t), power 2 (explicit one-sided
 *     d/dt of the Laplacian; the reference's form is implicit), power None
 *     (p+ <- alpha p+)                                           devito.py:1536-1548, 1602-1606
 *   - sponge boundary term vp^2 (2 damp p_t + damp^2 p)         boundaries/devito.py:37-69
 *   - CFS-PML auxiliary fields u3 -> u2 -> u1 per axis          boundaries/devito.py:72-193
 *   - source injection into p[t+1] after the stencil, receiver
 *     sampling of p[t+1], guarded every-f-th-step save          devito.py:357-364, 413-418
 *   - "compat" adjoint: time-reversed stencil, adjoint source injected into
 *     p_a[t-1] with -rec[t] dt^2 vp^2 (fw3d_mode), gradient
 *     accumulated on guarded steps over the interior             devito.py:745-761, 985-1024,
 *                                                                1115-1150, 916-920
 *   - "exact" adjoint: the transpose of the forward time-stepping map, run on
 *     mu = G*lambda (G = dt^2 vp^2 rho/(1+a)), which obeys the same backward
 *     stencil; it accumulates the pointwise-coefficient sensitivities that the
 *     host turns into d/dvp, d/drho, d/dalpha with the chain rule.
 *
 * Arrays handed in are unpadded, C order, extended grid shape n[0..dim-1],
 * last axis fastest. 2D runs as 3D with a leading axis of size 1.
 * REAL is float (oracle_*_f32) or double (oracle_*_f64).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>

#ifndef REAL
#define REAL float
#define SFX(name) name##_f32
#else
#define SFX(name) name##_f64
#endif

#define H 10 /* halo cells on each active axis (Devito uses space_order = 10) */

/* Planted mistakes, compiled only into the separate mutation build (-DORACLE_MUTATIONS,
 * _build/liboracle_mut_f64.so) that the oracle's physics checks load to show that each
 * check fails when the term it pins is wrong. The regular builds have no such switch. */
#ifdef ORACLE_MUTATIONS
static int oracle_mutation = 0;
void SFX(oracle_set_mutation)(int m) { oracle_mutation = m; }
#define MUT(k) (oracle_mutation == (k))
#else
#define MUT(k) 0
#endif
enum {
    MUT_ATT0_TWICE = 1,    /* attenuation power 0: 2 x the damping coefficient alpha vp */
    MUT_ATT0_SIGN = 2,     /* attenuation power 0: damping coefficient of the wrong sign */
    MUT_ATT2_TWICE = 3,    /* attenuation power 2: 2 x Cbeta */
    MUT_ATT2_SIGN = 4,     /* attenuation power 2: Cbeta of the wrong sign */
    MUT_SPONGE_TWICE = 5,  /* sponge: 2 x the vp^2 damp p_t coefficient */
    MUT_SPONGE_SIGN = 6,   /* sponge: vp^2 damp p_t coefficient of the wrong sign */
    MUT_PML_DSIGMA = 7,    /* CFS-PML: sigma' of the wrong sign in the forward aux updates */
    MUT_PML_D2 = 8,        /* CFS-PML: 2 x the sigma d2p term of u2 */
    MUT_PML_ALPHA = 9,     /* CFS-PML: alpha of the wrong sign in (alpha + sigma) */
    MUT_PML_ADJ_FORM = 10, /* CFS-PML: the adjoint run uses the forward aux form */
    MUT_PML_U1 = 11,       /* CFS-PML: boundary term sum u1 of the wrong sign */
};

typedef struct {
    int dim;              /* 2 or 3 */
    int n[3];             /* extended shape */
    double h[3];          /* spacing */
    int nabs[3];          /* absorbing layer thickness per axis */
    double dt;
    int nt;
    int kernel;           /* 2: OT2, 4: OT4 */
    int rho_mode;         /* 0 none, 1 constant (div grad), 2 variable */
    int att_power;        /* -2 no attenuation, 0, 2, -1 (power None) */
    int boundary;         /* -1 none, 0 sponge, 1 CFS-PML */
    int diff_source;
    int undersampling;    /* f */
    int save_raw;         /* 0: save p.dt2, 1: save p */
    int tb0, tb1;         /* time bounds */
    int grad_vp, grad_rho, grad_src;
} oracle_cfg;

/* SO10 weights (Fornberg closed forms) */
static const double W2[6] = {-5269.0 / 1800.0, 5.0 / 3.0, -5.0 / 21.0, 5.0 / 126.0, -5.0 / 1008.0, 1.0 / 3150.0};
static const double W1[6] = {0.0, 5.0 / 6.0, -5.0 / 21.0, 5.0 / 84.0, -5.0 / 504.0, 1.0 / 1260.0};
static const double WS[5] = {19845.0 / 16384.0, -735.0 / 8192.0, 567.0 / 40960.0, -405.0 / 229376.0,
                              35.0 / 294912.0};

typedef struct {
    const oracle_cfg *c;
    int act[3];           /* active flag per internal axis */
    int n[3];             /* internal shape (3 axes) */
    int N[3];             /* padded internal shape */
    long s[3];            /* padded strides */
    long total;           /* padded size */
    long npts;            /* unpadded size */
    int nabs[3];
    REAL ih[3], ih2[3];
    REAL w2[3][6], w1[3][6], ws[3][5];
} geom;

static void geom_init(geom *g, const oracle_cfg *c) {
    g->c = c;
    int off = 3 - c->dim;
    for (int a = 0; a < 3; ++a) {
        int d = a - off;
        g->act[a] = d >= 0;
        g->n[a] = d >= 0 ? c->n[d] : 1;
        g->nabs[a] = d >= 0 ? c->nabs[d] : 0;
        g->N[a] = g->n[a] + (g->act[a] ? 2 * H : 0);
        double h = d >= 0 ? c->h[d] : 1.0;
        g->ih[a] = (REAL)(1.0 / h);
        g->ih2[a] = (REAL)(1.0 / (h * h));
        for (int k = 0; k < 6; ++k) {
            g->w2[a][k] = (REAL)(W2[k] / (h * h));
            g->w1[a][k] = (REAL)(W1[k] / h);
        }
        for (int k = 0; k < 5; ++k) g->ws[a][k] = (REAL)(WS[k] / h);
    }
    g->s[2] = 1;
    g->s[1] = g->N[2];
    g->s[0] = (long)g->N[1] * g->N[2];
    g->total = (long)g->N[0] * g->N[1] * g->N[2];
    g->npts = (long)g->n[0] * g->n[1] * g->n[2];
}

static inline long pidx(const geom *g, int i, int j, int k) {
    return (long)(i + (g->act[0] ? H : 0)) * g->s[0] + (long)(j + (g->act[1] ? H : 0)) * g->s[1] +
           (k + (g->act[2] ? H : 0));
}

static REAL *zalloc(long n) { return (REAL *)calloc((size_t)n, sizeof(REAL)); }

/* copy unpadded float array into padded REAL array; edge=1 replicates edges into the halo */
static void pad_in(const geom *g, REAL *dst, const float *src, int edge) {
    for (int i = -(g->act[0] ? H : 0); i < g->n[0] + (g->act[0] ? H : 0); ++i)
        for (int j = -(g->act[1] ? H : 0); j < g->n[1] + (g->act[1] ? H : 0); ++j)
            for (int k = -(g->act[2] ? H : 0); k < g->n[2] + (g->act[2] ? H : 0); ++k) {
                int ii = i, jj = j, kk = k;
                int inside = i >= 0 && i < g->n[0] && j >= 0 && j < g->n[1] && k >= 0 && k < g->n[2];
                if (!inside && !edge) continue;
                ii = ii < 0 ? 0 : (ii >= g->n[0] ? g->n[0] - 1 : ii);
                jj = jj < 0 ? 0 : (jj >= g->n[1] ? g->n[1] - 1 : jj);
                kk = kk < 0 ? 0 : (kk >= g->n[2] ? g->n[2] - 1 : kk);
                dst[pidx(g, i, j, k)] = (REAL)src[((long)ii * g->n[1] + jj) * g->n[2] + kk];
            }
}

static void unpad_out(const geom *g, float *dst, const REAL *src) {
    for (int i = 0; i < g->n[0]; ++i)
        for (int j = 0; j < g->n[1]; ++j)
            for (int k = 0; k < g->n[2]; ++k)
                dst[((long)i * g->n[1] + j) * g->n[2] + k] = (float)src[pidx(g, i, j, k)];
}

static inline int in_interior(const geom *g, int i, int j, int k) {
    return i >= g->nabs[0] && i < g->n[0] - g->nabs[0] && j >= g->nabs[1] && j < g->n[1] - g->nabs[1] &&
           k >= g->nabs[2] && k < g->n[2] - g->nabs[2];
}

#define LOOP_DOMAIN                                         \
    _Pragma("omp parallel for collapse(2) schedule(static)") \
    for (int i = 0; i < g->n[0]; ++i)                       \
        for (int j = 0; j < g->n[1]; ++j)                   \
            for (int k = 0; k < g->n[2]; ++k)

/* standard SO10 Laplacian of u at padded index p */
static inline REAL lap(const geom *g, const REAL *u, long p) {
    REAL acc = 0;
    for (int a = 0; a < 3; ++a) {
        if (!g->act[a]) continue;
        long s = g->s[a];
        REAL v = g->w2[a][0] * u[p];
        for (int k = 1; k <= 5; ++k) v += g->w2[a][k] * (u[p + k * s] + u[p - k * s]);
        acc += v;
    }
    return acc;
}

static inline REAL d1(const geom *g, const REAL *u, long p, int a) {
    long s = g->s[a];
    REAL v = 0;
    for (int k = 1; k <= 5; ++k) v += g->w1[a][k] * (u[p + k * s] - u[p - k * s]);
    return v;
}

static inline REAL d2(const geom *g, const REAL *u, long p, int a) {
    long s = g->s[a];
    REAL v = g->w2[a][0] * u[p];
    for (int k = 1; k <= 5; ++k) v += g->w2[a][k] * (u[p + k * s] + u[p - k * s]);
    return v;
}

/* staggered derivative at q+1/2 */
static inline REAL dplus(const geom *g, const REAL *u, long q, int a) {
    long s = g->s[a];
    REAL v = 0;
    for (int k = 0; k < 5; ++k) v += g->ws[a][k] * (u[q + (k + 1) * s] - u[q - k * s]);
    return v;
}

/* staggered derivative at node p of a half-point array (gq[q] is the value at q+1/2) */
static inline REAL dminus(const geom *g, const REAL *gq, long p, int a) {
    long s = g->s[a];
    REAL v = 0;
    for (int k = 0; k < 5; ++k) v += g->ws[a][k] * (gq[p + k * s] - gq[p - (k + 1) * s]);
    return v;
}

/* work buffers for the div(b grad) operator */
typedef struct {
    REAL *stag[3];  /* half-point gradients per axis */
    REAL *bbar[3];  /* buoyancy averaged to half points (NULL: constant density) */
} kops;

/* K u: Laplacian (rho none) or sum_a D-(bbar D+ u) (rho const / var); out over the domain */
static void apply_K(const geom *g, const kops *ko, int rho_mode, const REAL *u, REAL *out) {
    if (rho_mode == 0) {
        LOOP_DOMAIN { long p = pidx(g, i, j, k); out[p] = lap(g, u, p); }
        return;
    }
    for (int a = 0; a < 3; ++a) {
        if (!g->act[a]) continue;
        REAL *st = ko->stag[a];
        const REAL *bb = ko->bbar[a];
        int lo[3] = {0, 0, 0}, hi[3];
        for (int b = 0; b < 3; ++b) hi[b] = g->n[b];
        lo[a] = -5; hi[a] = g->n[a] + 5;
#pragma omp parallel for collapse(2) schedule(static)
        for (int i = lo[0]; i < hi[0]; ++i)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int k = lo[2]; k < hi[2]; ++k) {
                    long q = pidx(g, i, j, k);
                    REAL v = dplus(g, u, q, a);
                    st[q] = bb ? bb[q] * v : v;
                }
    }
    LOOP_DOMAIN {
        long p = pidx(g, i, j, k);
        REAL acc = 0;
        for (int a = 0; a < 3; ++a)
            if (g->act[a]) acc += dminus(g, ko->stag[a], p, a);
        out[p] = acc;
    }
}

/* medium and per-point step coefficients */
typedef struct {
    REAL *vp, *m, *rho, *b, *alpha, *damp; /* padded */
    REAL *A, *B, *G, *Cq, *Cbeta;         /* padded, domain-valid */
    REAL *pml[3][4];                       /* per internal axis: sigma, alpha, dsigma, dalpha (length n[a]) */
} medium;

static void medium_init(const geom *g, medium *md, const float *vp, const float *rho, const float *alpha,
                        const float *damp, const float *pml, int pml_stride) {
    const oracle_cfg *c = g->c;
    memset(md, 0, sizeof(*md));
    md->vp = zalloc(g->total);
    pad_in(g, md->vp, vp, 1);
    if (c->rho_mode == 2) {
        md->rho = zalloc(g->total);
        md->b = zalloc(g->total);
        pad_in(g, md->rho, rho, 1);
        for (long p = 0; p < g->total; ++p) md->b[p] = md->rho[p] != 0 ? (REAL)1 / md->rho[p] : 0;
    }
    if (c->att_power != -2) {
        md->alpha = zalloc(g->total);
        pad_in(g, md->alpha, alpha, 1);
    }
    if (c->boundary == 0) {
        md->damp = zalloc(g->total);
        pad_in(g, md->damp, damp, 0);
    }
    if (c->boundary == 1) {
        int off = 3 - c->dim;
        for (int d = 0; d < c->dim; ++d)
            for (int r = 0; r < 4; ++r) {
                md->pml[d + off][r] = zalloc(g->n[d + off]);
                for (int x = 0; x < g->n[d + off]; ++x)
                    md->pml[d + off][r][x] = (REAL)pml[((long)d * 4 + r) * pml_stride + x];
            }
    }
    md->m = zalloc(g->total);
    md->A = zalloc(g->total);
    md->B = zalloc(g->total);
    md->G = zalloc(g->total);
    md->Cq = zalloc(g->total);
    if (c->att_power == 2) md->Cbeta = zalloc(g->total);
    const REAL dt = (REAL)c->dt, dt2 = (REAL)(c->dt * c->dt);
    LOOP_DOMAIN {
        long p = pidx(g, i, j, k);
        REAL v = md->vp[p];
        REAL v2 = v * v;
        REAL m = c->rho_mode == 2 ? v2 * md->rho[p] : v2;
        REAL a = 0, e = 0, W = 1;
        if (c->att_power == 0) {
            REAL q = dt * md->alpha[p] * v;
            if (MUT(MUT_ATT0_TWICE)) q *= 2;
            if (MUT(MUT_ATT0_SIGN)) q = -q;
            a += q;
        }
        if (c->boundary == 0) {
            REAL D = md->damp[p];
            REAL q = dt * v2 * D;
            if (MUT(MUT_SPONGE_TWICE)) q *= 2;
            if (MUT(MUT_SPONGE_SIGN)) q = -q;
            a += q;
            e = dt2 * v2 * D * D;
        }
        if (c->att_power == -1) W = md->alpha[p];
        REAL ia = (REAL)1 / ((REAL)1 + a);
        md->m[p] = m;
        md->A[p] = W * ((REAL)2 - e) * ia;
        md->B[p] = -W * ((REAL)1 - a) * ia;
        md->Cq[p] = W * dt2 * ia;
        md->G[p] = md->Cq[p] * m;
        if (c->att_power == 2) {
            md->Cbeta[p] = md->Cq[p] * (REAL)2 * md->alpha[p] * v2 * v / dt;
            if (MUT(MUT_ATT2_TWICE)) md->Cbeta[p] *= 2;
            if (MUT(MUT_ATT2_SIGN)) md->Cbeta[p] = -md->Cbeta[p];
        }
    }
}

static void medium_free(medium *md) {
    free(md->vp); free(md->m); free(md->rho); free(md->b); free(md->alpha); free(md->damp);
    free(md->A); free(md->B); free(md->G); free(md->Cq); free(md->Cbeta);
    for (int a = 0; a < 3; ++a)
        for (int r = 0; r < 4; ++r) free(md->pml[a][r]);
}

static void kops_init(const geom *g, const medium *md, int rho_mode, kops *ko) {
    memset(ko, 0, sizeof(*ko));
    if (rho_mode == 0) return;
    for (int a = 0; a < 3; ++a) {
        if (!g->act[a]) continue;
        ko->stag[a] = zalloc(g->total);
        if (rho_mode == 2) {
            ko->bbar[a] = zalloc(g->total);
            long s = g->s[a];
            /* half point q+1/2 along a: arithmetic mean of the edge-padded buoyancy */
            for (long q = 0; q + s < g->total; ++q) ko->bbar[a][q] = (REAL)0.5 * (md->b[q] + md->b[q + s]);
        }
    }
}

static void kops_free(kops *ko) {
    for (int a = 0; a < 3; ++a) { free(ko->stag[a]); free(ko->bbar[a]); }
}

/* PML auxiliary fields */
typedef struct { REAL *u1[3], *u2[3], *u3[3]; REAL *n1[3], *n2[3], *n3[3]; } pmlaux;

static void pml_alloc(const geom *g, pmlaux *px) {
    memset(px, 0, sizeof(*px));
    for (int a = 0; a < 3; ++a) {
        if (!g->act[a]) continue;
        px->u1[a] = zalloc(g->total); px->u2[a] = zalloc(g->total); px->u3[a] = zalloc(g->total);
        px->n1[a] = zalloc(g->total); px->n2[a] = zalloc(g->total); px->n3[a] = zalloc(g->total);
    }
}

static void pml_free(pmlaux *px) {
    for (int a = 0; a < 3; ++a) {
        free(px->u1[a]); free(px->u2[a]); free(px->u3[a]);
        free(px->n1[a]); free(px->n2[a]); free(px->n3[a]);
    }
}

/* u3 -> u2 -> u1 updates on the left/right slabs of each axis (boundaries/devito.py:133-171).
 * backward=1 applies the as-written transposed form (.dt.T, u_di.T). */
static void pml_step(const geom *g, const medium *md, pmlaux *px, const REAL *F, int backward) {
    const REAL dt = (REAL)g->c->dt;
    for (int a = 0; a < 3; ++a) {
        if (!g->act[a]) continue;
        const REAL *sg = md->pml[a][0], *al = md->pml[a][1], *dsg = md->pml[a][2], *dal = md->pml[a][3];
        int na = g->nabs[a];
        REAL *u1 = px->u1[a], *u2 = px->u2[a], *u3 = px->u3[a];
        REAL *n1 = px->n1[a], *n2 = px->n2[a], *n3 = px->n3[a];
        LOOP_DOMAIN {
            int x = a == 0 ? i : (a == 1 ? j : k);
            if (!(x < na || x >= g->n[a] - na)) continue;
            long p = pidx(g, i, j, k);
            REAL s = sg[x], al_ = al[x], ds = dsg[x], da = dal[x];
            if (MUT(MUT_PML_DSIGMA) && !backward) ds = -ds;
            if (MUT(MUT_PML_ALPHA)) al_ = -al_;
            REAL dF = d1(g, F, p, a), d2F = d2(g, F, p, a);
            REAL v3, v2, v1;
            if (!backward || MUT(MUT_PML_ADJ_FORM)) {
                v3 = u3[p] - dt * ((al_ + s) * u3[p] - s * s * (ds + da) * dF);
                v2 = u2[p] - dt * ((al_ + s) * u2[p] - s * ((REAL)2 * ds * dF + da * dF +
                                                               (MUT(MUT_PML_D2) ? (REAL)2 : (REAL)1) * s * d2F) + v3);
                v1 = u1[p] - dt * ((al_ + s) * u1[p] - (ds * dF + (REAL)2 * s * d2F) + v2);
            } else {
                v3 = u3[p] - dt * ((al_ + s) * u3[p] + s * s * (ds + da) * dF);
                v2 = u2[p] - dt * ((al_ + s) * u2[p] + s * ((REAL)2 * ds * dF + da * dF) - s * s * d2F + v3);
                v1 = u1[p] - dt * ((al_ + s) * u1[p] + ds * dF - (REAL)2 * s * d2F + v2);
            }
            n3[p] = v3; n2[p] = v2; n1[p] = v1;
        }
        /* commit (time_order 1: the new level replaces the old) */
        REAL *t;
        t = px->u1[a]; px->u1[a] = px->n1[a]; px->n1[a] = t;
        t = px->u2[a]; px->u2[a] = px->n2[a]; px->n2[a] = t;
        t = px->u3[a]; px->u3[a] = px->n3[a]; px->n3[a] = t;
    }
}

/* One time step of the state or (compat/exact) adjoint stencil:
 * out = A u + B other + Cq*(Q(u) + extra), Q = m K u (OT2) or L2 + dt^2/12 m K L2 (OT4).
 * 'other' is p[t-1] forward and p[t+1] backward; 'lap_prev' carries Lap(other) for power-2
 * attenuation. scratch: Ku, L2, KL2 work arrays. */
typedef struct { REAL *Ku, *L2, *KL2, *lapu, *lapo; } work;

static void step(const geom *g, const medium *md, const kops *ko, pmlaux *px, work *wk,
                 const REAL *u, const REAL *other, REAL *out, int backward) {
    const oracle_cfg *c = g->c;
    const REAL dt2 = (REAL)(c->dt * c->dt);
    apply_K(g, ko, c->rho_mode, u, wk->Ku);
    const REAL *F = u;
    if (c->kernel == 4) {
        LOOP_DOMAIN { long p = pidx(g, i, j, k); wk->L2[p] = md->m[p] * wk->Ku[p]; }
        apply_K(g, ko, c->rho_mode, wk->L2, wk->KL2);
        F = wk->L2;
    }
    if (c->boundary == 1) {
        /* with OT4 the aux fields read the laplacian Function before it is
         * refreshed in this step (program order of eq_before vs stencils) */
        pml_step(g, md, px, c->kernel == 4 ? wk->lapo : F, backward);
    }
    if (c->att_power == 2) {
        LOOP_DOMAIN {
            long p = pidx(g, i, j, k);
            wk->lapu[p] = lap(g, u, p);
        }
    }
    LOOP_DOMAIN {
        long p = pidx(g, i, j, k);
        REAL Q = md->m[p] * wk->Ku[p];
        if (c->kernel == 4) Q = wk->L2[p] + dt2 / (REAL)12 * md->m[p] * wk->KL2[p];
        REAL acc = md->Cq[p] * Q;
        if (c->att_power == 2) acc += md->Cbeta[p] * (wk->lapu[p] - lap(g, other, p));
        if (c->boundary == 1) {
            REAL s1 = 0;
            for (int a = 0; a < 3; ++a)
                if (g->act[a]) s1 += px->u1[a][p];
            if (MUT(MUT_PML_U1)) s1 = -s1;
            acc -= md->Cq[p] * md->vp[p] * md->vp[p] * s1;
        }
        out[p] = md->A[p] * u[p] + md->B[p] * other[p] + acc;
    }
    if (c->kernel == 4 && c->boundary == 1) memcpy(wk->lapo, wk->L2, sizeof(REAL) * g->total);
}

/* sparse taps: idx[n*ntap] are unpadded linear indices (-1 = dropped) */
static inline long tap_pidx(const geom *g, long lin) {
    int k = (int)(lin % g->n[2]);
    long r = lin / g->n[2];
    int j = (int)(r % g->n[1]);
    int i = (int)(r / g->n[1]);
    return pidx(g, i, j, k);
}

static void inject(const geom *g, REAL *field, int npt, int ntap, const int64_t *idx, const REAL *coef,
                   const float *values /* per point */, REAL sign) {
    for (int s = 0; s < npt; ++s) {
        REAL v = (REAL)values[s] * sign;
        for (int q = 0; q < ntap; ++q) {
            long lin = idx[(long)s * ntap + q];
            if (lin < 0) continue;
            field[tap_pidx(g, lin)] += coef[(long)s * ntap + q] * v;
        }
    }
}

static void gather(const geom *g, const REAL *field, int npt, int ntap, const int64_t *idx, const float *w,
                   float *out) {
    for (int r = 0; r < npt; ++r) {
        REAL acc = 0;
        for (int q = 0; q < ntap; ++q) {
            long lin = idx[(long)r * ntap + q];
            if (lin < 0) continue;
            acc += (REAL)w[(long)r * ntap + q] * field[tap_pidx(g, lin)];
        }
        out[r] = (float)acc;
    }
}

static REAL *tap_coef(const geom *g, const medium *md, int npt, int ntap, const int64_t *idx, const float *w,
                      int kind /* 0: source injection, 1: compat adjoint, 2: exact adjoint (G) */) {
    const oracle_cfg *c = g->c;
    double hmax = 0;
    for (int d = 0; d < c->dim; ++d) hmax = c->h[d] > hmax ? c->h[d] : hmax;
    REAL *out = zalloc((long)npt * ntap);
    const REAL dt = (REAL)c->dt;
    REAL scale = c->diff_source ? (REAL)(2 * c->dt * c->dt / hmax) : (REAL)(2 * c->dt / hmax);
    for (long q = 0; q < (long)npt * ntap; ++q) {
        if (idx[q] < 0) continue;
        long p = tap_pidx(g, idx[q]);
        REAL v = md->vp[p];
        if (kind == 0) out[q] = (REAL)w[q] * scale * v;
        else if (kind == 1) out[q] = (REAL)w[q] * dt * dt * v * v;
        else out[q] = (REAL)w[q] * md->G[p];
    }
    return out;
}

/* ------------------------------------------------------------------------------------------ */

/* Forward run (run_forward, devito.py:515-581): t = 1 .. tb1-2.
 * rec_out (nt, nrec) time-major; saved (nsave, npts) when nsave > 0 (full-grid snapshots);
 * final (2, npts) optional: p[t_last+1], p[t_last]. */
int SFX(oracle_forward)(const oracle_cfg *c, const float *vp, const float *rho, const float *alpha,
                        const float *damp, const float *pml, int pml_stride,
                        int nsrc, int src_ntap, const int64_t *src_idx, const float *src_w, const float *src,
                        int nrec, int rec_ntap, const int64_t *rec_idx, const float *rec_w, float *rec_out,
                        float *saved, int nsave, float *final_state) {
    geom G0, *g = &G0;
    geom_init(g, c);
    medium md;
    medium_init(g, &md, vp, rho, alpha, damp, pml, pml_stride);
    kops ko;
    kops_init(g, &md, c->rho_mode, &ko);
    pmlaux px;
    if (c->boundary == 1) pml_alloc(g, &px);
    work wk = {zalloc(g->total), c->kernel == 4 ? zalloc(g->total) : NULL,
               c->kernel == 4 ? zalloc(g->total) : NULL, zalloc(g->total), zalloc(g->total)};
    REAL *src_c = tap_coef(g, &md, nsrc, src_ntap, src_idx, src_w, 0);
    REAL *buf[3] = {zalloc(g->total), zalloc(g->total), zalloc(g->total)};
    memset(rec_out, 0, sizeof(float) * (size_t)c->nt * nrec);
    const int f = c->undersampling > 0 ? c->undersampling : 1;
    const REAL idt2 = (REAL)(1.0 / (c->dt * c->dt));
    int tlast = c->tb1 - 2;
    for (int t = 1; t <= tlast; ++t) {
        REAL *pm = buf[(t + 2) % 3], *p0 = buf[t % 3], *pp = buf[(t + 1) % 3];
        step(g, &md, &ko, &px, &wk, p0, pm, pp, 0);
        inject(g, pp, nsrc, src_ntap, src_idx, src_c, src + (long)t * nsrc, 1);
        gather(g, pp, nrec, rec_ntap, rec_idx, rec_w, rec_out + (long)t * nrec);
        if (nsave > 0 && t % f == 0 && t >= c->tb0 && t <= c->tb1) {
            int slot = t / f;
            if (slot < nsave) {
                float *dst = saved + (long)slot * g->npts;
                LOOP_DOMAIN {
                    long p = pidx(g, i, j, k);
                    REAL v = c->save_raw ? p0[p] : (pp[p] - (REAL)2 * p0[p] + pm[p]) * idt2;
                    dst[((long)i * g->n[1] + j) * g->n[2] + k] = (float)v;
                }
            }
        }
    }
    if (final_state) {
        unpad_out(g, final_state, buf[(tlast + 1) % 3]);
        unpad_out(g, final_state + g->npts, buf[tlast % 3]);
    }
    free(buf[0]); free(buf[1]); free(buf[2]); free(src_c);
    free(wk.Ku); free(wk.L2); free(wk.KL2); free(wk.lapu); free(wk.lapo);
    if (c->boundary == 1) pml_free(&px);
    kops_free(&ko);
    medium_free(&md);
    return 0;
}

/* staggered derivative at node i evaluated at i-1/2 (grad(f, shift=-0.5)) */
static inline REAL dminus_node(const geom *g, const REAL *u, long p, int a) {
    long s = g->s[a];
    REAL v = 0;
    for (int k = 0; k < 5; ++k) v += g->ws[a][k] * (u[p + k * s] - u[p - (k + 1) * s]);
    return v;
}

/* Compat adjoint (run_adjoint, devito.py:875-920): t = tb1-1-f down to tb0+1.
 * res (nt, nrec); saved (nsave, npts). Outputs are full extended grid (npts); the
 * gradient is accumulated on the interior only. grad_src (nt, nsrc).
 * final_state (2, npts): p_a[t_end-1], p_a[t_end]. */
int SFX(oracle_adjoint_compat)(const oracle_cfg *c, const float *vp, const float *rho, const float *alpha,
                               const float *damp, const float *pml, int pml_stride,
                               int nsrc, int src_ntap, const int64_t *src_idx, const float *src_w,
                               int nrec, int rec_ntap, const int64_t *rec_idx, const float *rec_w, const float *res,
                               const float *saved, int nsave,
                               float *grad_vp, float *prec_vp, float *grad_rho, float *prec_rho,
                               float *grad_src, float *final_state) {
    geom G0, *g = &G0;
    geom_init(g, c);
    medium md;
    medium_init(g, &md, vp, rho, alpha, damp, pml, pml_stride);
    kops ko;
    kops_init(g, &md, c->rho_mode, &ko);
    pmlaux px;
    if (c->boundary == 1) pml_alloc(g, &px);
    work wk = {zalloc(g->total), c->kernel == 4 ? zalloc(g->total) : NULL,
               c->kernel == 4 ? zalloc(g->total) : NULL, zalloc(g->total), zalloc(g->total)};
    REAL *rec_c = tap_coef(g, &md, nrec, rec_ntap, rec_idx, rec_w, 1);
    REAL *buf[3] = {zalloc(g->total), zalloc(g->total), zalloc(g->total)};
    REAL *ps = zalloc(g->total), *psm = NULL, *psp = NULL, *grf = NULL, *bpad = NULL;
    double *gv = calloc(g->npts, sizeof(double)), *pv = calloc(g->npts, sizeof(double));
    double *gr = NULL, *pr = NULL;
    if (c->grad_rho) {
        gr = calloc(g->npts, sizeof(double)); pr = calloc(g->npts, sizeof(double));
        grf = zalloc(g->total);
        bpad = zalloc(g->total);
        if (md.b) memcpy(bpad, md.b, sizeof(REAL) * g->total);
        else for (long p = 0; p < g->total; ++p) bpad[p] = (REAL)1e-3; /* unused without rho */
    }
    if (c->save_raw) { psm = zalloc(g->total); psp = zalloc(g->total); }
    const int f = c->undersampling > 0 ? c->undersampling : 1;
    if (grad_src) memset(grad_src, 0, sizeof(float) * (size_t)c->nt * nsrc);
    const REAL fdt2 = (REAL)(1.0 / ((double)f * c->dt * (double)f * c->dt));
    int tstart = c->tb1 - 1 - f, tend = c->tb0 + 1;
    /* p_a[t+1] = buf[(t+1)%3], p_a[t] = buf[t%3], p_a[t-1] = buf[(t+2)%3] */
    for (int t = tstart; t >= tend; --t) {
        REAL *pp = buf[(t + 1) % 3], *p0 = buf[t % 3], *pm = buf[(t + 2) % 3];
        step(g, &md, &ko, &px, &wk, p0, pp, pm, 1);
        inject(g, pm, nrec, rec_ntap, rec_idx, rec_c, res + (long)t * nrec, -1);
        if (grad_src) {
            float *row = grad_src + (long)t * nsrc;
            gather(g, p0, nsrc, src_ntap, src_idx, src_w, row);
            for (int s = 0; s < nsrc; ++s) row[s] = -row[s];
        }
        if (t % f == 0 && t >= c->tb0 && t <= c->tb1 && nsave > 0) {
            int slot = t / f;
            if (slot >= nsave) continue;
            pad_in(g, ps, saved + (long)slot * g->npts, 0);
            if (c->save_raw) {
                if (slot + 1 < nsave) pad_in(g, psp, saved + (long)(slot + 1) * g->npts, 0);
                else memset(psp, 0, sizeof(REAL) * g->total);
                if (slot >= 1) pad_in(g, psm, saved + (long)(slot - 1) * g->npts, 0);
                else memset(psm, 0, sizeof(REAL) * g->total);
            }
            if (c->grad_rho) {
                LOOP_DOMAIN {
                    if (!in_interior(g, i, j, k)) continue;
                    long p = pidx(g, i, j, k);
                    REAL acc = 0;
                    for (int a = 0; a < 3; ++a)
                        if (g->act[a]) acc += dminus_node(g, bpad, p, a) * dminus_node(g, ps, p, a);
                    grf[p] = -acc - bpad[p] * lap(g, ps, p);
                }
            }
#pragma omp parallel for collapse(2) schedule(static)
            for (int i = 0; i < g->n[0]; ++i)
                for (int j = 0; j < g->n[1]; ++j)
                    for (int k = 0; k < g->n[2]; ++k) {
                        if (!in_interior(g, i, j, k)) continue;
                        long p = pidx(g, i, j, k);
                        long q = ((long)i * g->n[1] + j) * g->n[2] + k;
                        if (c->grad_vp) {
                            REAL pdt = c->save_raw ? (psp[p] - (REAL)2 * ps[p] + psm[p]) * fdt2 : ps[p];
                            gv[q] += (double)(pdt * p0[p]);
                            pv[q] += (double)(pdt * pdt);
                        }
                        if (c->grad_rho) {
                            gr[q] += (double)(grf[p] * p0[p]);
                            pr[q] += (double)(grf[p] * grf[p]);
                        }
                    }
        }
    }
    if (grad_vp) for (long q = 0; q < g->npts; ++q) grad_vp[q] = (float)gv[q];
    if (prec_vp) for (long q = 0; q < g->npts; ++q) prec_vp[q] = (float)pv[q];
    if (grad_rho && gr) for (long q = 0; q < g->npts; ++q) grad_rho[q] = (float)gr[q];
    if (prec_rho && pr) for (long q = 0; q < g->npts; ++q) prec_rho[q] = (float)pr[q];
    if (final_state) {
        unpad_out(g, final_state, buf[(tend + 2) % 3]);
        unpad_out(g, final_state + g->npts, buf[tend % 3]);
    }
    free(buf[0]); free(buf[1]); free(buf[2]); free(rec_c); free(ps); free(psm); free(psp); free(grf); free(bpad);
    free(gv); free(pv); free(gr); free(pr);
    free(wk.Ku); free(wk.L2); free(wk.KL2); free(wk.lapu); free(wk.lapo);
    if (c->boundary == 1) pml_free(&px);
    kops_free(&ko);
    medium_free(&md);
    return 0;
}

/* Exact adjoint of the forward map for J = sum_t <res[t], rec[t]>.
 * states (nt, npts): every forward level u_0 .. u_{nt-1} (raw, full grid).
 * Outputs (npts each, full grid), accumulated over every equation t = 1..nt-2:
 *   gA = sum lambda_{t+1} u_t,  gB = sum lambda_{t+1} u_{t-1},
 *   gG = sum lambda_{t+1} (K u_t [+ dt^2/12 K m K u_t])      (OT4 adds gm = dt^2/12 sum (K mu)(K u))
 *   gb = d/db through the half-point buoyancy (rho variable)
 * lam_src (nt, nsrc*src_ntap): lambda_{t+1} at each source tap for step t.
 * final_state (2, npts): mu_2, mu_3. Supports OT2 (all rho modes) and OT4 (rho none/const);
 * sponge or no boundary; attenuation none / power 0 / power None. */
int SFX(oracle_adjoint_exact)(const oracle_cfg *c, const float *vp, const float *rho, const float *alpha,
                              const float *damp,
                              int nsrc, int src_ntap, const int64_t *src_idx,
                              int nrec, int rec_ntap, const int64_t *rec_idx, const float *rec_w, const float *res,
                              const float *states,
                              float *gA, float *gB, float *gG, float *gm, float *gb, float *lam_src,
                              float *final_state) {
    if (c->boundary == 1 || c->att_power == 2 || (c->kernel == 4 && c->rho_mode == 2)) return -1;
    geom G0, *g = &G0;
    geom_init(g, c);
    medium md;
    medium_init(g, &md, vp, rho, alpha, damp, NULL, 0);
    kops ko;
    kops_init(g, &md, c->rho_mode, &ko);
    work wk = {zalloc(g->total), c->kernel == 4 ? zalloc(g->total) : NULL,
               c->kernel == 4 ? zalloc(g->total) : NULL, zalloc(g->total), zalloc(g->total)};
    REAL *rec_c = tap_coef(g, &md, nrec, rec_ntap, rec_idx, rec_w, 2);
    REAL *buf[3] = {zalloc(g->total), zalloc(g->total), zalloc(g->total)};
    REAL *u = zalloc(g->total), *um = zalloc(g->total), *Ku = zalloc(g->total), *Kmu = NULL, *tmp = NULL;
    REAL *stg_u[3] = {0}, *stg_m[3] = {0};
    if (c->kernel == 4) { Kmu = zalloc(g->total); tmp = zalloc(g->total); }
    if (c->rho_mode == 2)
        for (int a = 0; a < 3; ++a)
            if (g->act[a]) { stg_u[a] = zalloc(g->total); stg_m[a] = zalloc(g->total); }
    double *aA = calloc(g->npts, sizeof(double)), *aB = calloc(g->npts, sizeof(double));
    double *aG = calloc(g->npts, sizeof(double)), *am = calloc(g->npts, sizeof(double));
    double *ab = calloc(g->npts, sizeof(double));
    const int nt = c->nt;
    const REAL dt2 = (REAL)(c->dt * c->dt);
    if (lam_src) memset(lam_src, 0, sizeof(float) * (size_t)nt * nsrc * src_ntap);
    /* mu_k = buf[k%3], k = nt-1 .. 2 */
    for (int lv = nt - 1; lv >= 1; --lv) {
        REAL *mk = buf[lv % 3], *m1 = buf[(lv + 1) % 3], *m2 = buf[(lv + 2) % 3];
        if (lv >= 2) {
            step(g, &md, &ko, NULL, &wk, m1, m2, mk, 1);
            if (lv - 1 >= 1 && lv - 1 <= nt - 2)
                inject(g, mk, nrec, rec_ntap, rec_idx, rec_c, res + (long)(lv - 1) * nrec, 1);
        }
        /* equation t = lv uses mu_{lv+1} (m1), u_lv, u_{lv-1} */
        int t = lv;
        if (t < 1 || t > nt - 2) continue;
        pad_in(g, u, states + (long)t * g->npts, 0);
        pad_in(g, um, states + (long)(t - 1) * g->npts, 0);
        if (lam_src) {
            for (long q = 0; q < (long)nsrc * src_ntap; ++q) {
                if (src_idx[q] < 0) continue;
                long p = tap_pidx(g, src_idx[q]);
                lam_src[(long)t * nsrc * src_ntap + q] = (float)(m1[p] / md.G[p]);
            }
        }
        apply_K(g, &ko, c->rho_mode, u, Ku);
        if (c->kernel == 4) {
            /* gm += dt^2/12 (K mu)(K u); then Ku <- Ku + dt^2/12 K(m Ku) for gG */
            apply_K(g, &ko, c->rho_mode, m1, Kmu);
            LOOP_DOMAIN {
                long p = pidx(g, i, j, k);
                long q = ((long)i * g->n[1] + j) * g->n[2] + k;
                am[q] += (double)(dt2 / (REAL)12) * Kmu[p] * Ku[p];
                tmp[p] = md.m[p] * Ku[p];
            }
            apply_K(g, &ko, c->rho_mode, tmp, Kmu);
            LOOP_DOMAIN { long p = pidx(g, i, j, k); Ku[p] += dt2 / (REAL)12 * Kmu[p]; }
        }
        if (c->rho_mode == 2) {
            for (int a = 0; a < 3; ++a) {
                if (!g->act[a]) continue;
                int lo[3] = {0, 0, 0}, hi[3];
                for (int b = 0; b < 3; ++b) hi[b] = g->n[b];
                lo[a] = -5; hi[a] = g->n[a] + 5;
#pragma omp parallel for collapse(2) schedule(static)
                for (int i = lo[0]; i < hi[0]; ++i)
                    for (int j = lo[1]; j < hi[1]; ++j)
                        for (int kk = lo[2]; kk < hi[2]; ++kk) {
                            long q = pidx(g, i, j, kk);
                            stg_u[a][q] = dplus(g, u, q, a);
                            stg_m[a][q] = dplus(g, m1, q, a);
                        }
            }
        }
#pragma omp parallel for collapse(2) schedule(static)
        for (int i = 0; i < g->n[0]; ++i)
            for (int j = 0; j < g->n[1]; ++j)
                for (int kk = 0; kk < g->n[2]; ++kk) {
                    long p = pidx(g, i, j, kk);
                    long q = ((long)i * g->n[1] + j) * g->n[2] + kk;
                    REAL lam = m1[p] / md.G[p];
                    aA[q] += (double)lam * u[p];
                    aB[q] += (double)lam * um[p];
                    aG[q] += (double)lam * Ku[p];
                    if (c->rho_mode == 2) {
                        double acc = 0;
                        for (int a = 0; a < 3; ++a) {
                            if (!g->act[a]) continue;
                            long s = g->s[a];
                            acc += (double)stg_m[a][p] * stg_u[a][p] + (double)stg_m[a][p - s] * stg_u[a][p - s];
                        }
                        ab[q] += -0.5 * acc;
                    }
                }
    }
    for (long q = 0; q < g->npts; ++q) {
        if (gA) gA[q] = (float)aA[q];
        if (gB) gB[q] = (float)aB[q];
        if (gG) gG[q] = (float)aG[q];
        if (gm) gm[q] = (float)am[q];
        if (gb) gb[q] = (float)ab[q];
    }
    if (final_state) {
        unpad_out(g, final_state, buf[2 % 3]);
        unpad_out(g, final_state + g->npts, buf[3 % 3]);
    }
    free(buf[0]); free(buf[1]); free(buf[2]); free(rec_c); free(u); free(um); free(Ku); free(Kmu); free(tmp);
    for (int a = 0; a < 3; ++a) { free(stg_u[a]); free(stg_m[a]); }
    free(aA); free(aB); free(aG); free(am); free(ab);
    free(wk.Ku); free(wk.L2); free(wk.KL2); free(wk.lapu); free(wk.lapo);
    kops_free(&ko);
    medium_free(&md);
    return 0;
}

/* Spatial operator alone, for pinning the weights: out = K u on the domain, where K is
 * the SO10 Laplacian (rho_mode 0) or sum_d D-_d (bbar_d D+_d u) (rho_mode 1 / 2, with rho
 * giving the buoyancy b = 1/rho, edge-padded). u is zero outside the domain. */
int SFX(oracle_apply_K)(const oracle_cfg *c, const float *rho, const float *u, float *out) {
    geom G0, *g = &G0;
    geom_init(g, c);
    medium md;
    memset(&md, 0, sizeof(md));
    if (c->rho_mode == 2) {
        md.rho = zalloc(g->total);
        md.b = zalloc(g->total);
        pad_in(g, md.rho, rho, 1);
        for (long p = 0; p < g->total; ++p) md.b[p] = md.rho[p] != 0 ? (REAL)1 / md.rho[p] : 0;
    }
    kops ko;
    kops_init(g, &md, c->rho_mode, &ko);
    REAL *up = zalloc(g->total), *op = zalloc(g->total);
    pad_in(g, up, u, 0);
    apply_K(g, &ko, c->rho_mode, up, op);
    unpad_out(g, out, op);
    free(up); free(op); free(md.rho); free(md.b);
    kops_free(&ko);
    return 0;
}
