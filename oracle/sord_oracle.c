/*
 * ORACLE — test infrastructure only.
 *
 * A plain-C restatement of the reference SORD solver (serial process module,
 * OpenMP threading): it reads ``sord.in`` from the current directory and
 * writes the same output files the reference ``sord.x`` writes.  It is used
 * by the tests as the checker of the HIP solver and by bench.py as the timed
 * CPU baseline; the product never links, loads or runs it.
 *
 * Every routine follows the reference file and lines cited in its comment,
 * statement by statement and in the same floating-point evaluation order, so
 * that it can be pinned bitwise (or to the last ulp) against the amdflang
 * build of the reference (oracle/build_ref.sh).  The arithmetic of the
 * support operators (diff_nc, diff_cn, hourglass, node normals) is
 * transliterated from the reference on purpose: the oracle exists to state
 * that arithmetic exactly.
 *
 * Build: REALSIZE=4 (float) or REALSIZE=8 (double), see oracle/Makefile.
 *
 * Scope (what the reference itself supports without undefined behaviour):
 * rectangular / affine / expanded meshes, all boundary conditions, PML,
 * Kelvin-Voigt viscosity, hourglass control, slip-weakening rupture, and the
 * field-I/O ops '=', '+', '*', '.', '=<', '+<', '*<', '=>', '.>', the
 * cube ops '=@', '+@', '*@' on volume fields with the intended coordinates
 * (reference defect D9: parity there is with the solver, not the reference),
 * and '.<' with the intended buffer row (the reference reads it with a stale
 * loop index, sord/input_output.f90:350: parity there is with the solver and
 * the '.' op's closed form).  Not restated: random ops ('~', and gridnoise,
 * whose values depend on the Fortran RNG), cube ops on fault fields and
 * debug dumps.  Those abort with exit status 3.  The finite kinematic source
 * (sord/kinematic_source.f90) is restated.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef REALSIZE
#define REALSIZE 4
#endif
#if REALSIZE == 8
typedef double real;
#define HUGE_R 1.7976931348623157e308
#define R(x) (x)
#else
typedef float real;
#define HUGE_R 3.40282347e+38f
#define R(x) (x##f)
#endif

#if REALSIZE == 8
#define EXP exp
#define SQRT sqrt
#define COS cos
#define SIN sin
#define FABS fabs
#else
#define EXP expf
#define SQRT sqrtf
#define COS cosf
#define SIN sinf
#define FABS fabsf
#endif

static void die(int code, const char *msg) {
    fprintf(stderr, "%s\n", msg);
    exit(code);
}

/* ------------------------------------------------------------------------ */
/* globals: sord/globals.f90:5-107                                          */
/* ------------------------------------------------------------------------ */
static int bc1[3], bc2[3], debug_, faultopening, i1pml[3], i2pml[3], itio, itstats, mpin, mpout,
    n1expand[3], n2expand[3], npml, nproc3[3], nsource, nthread, ppml, shape_[4];
static real affine[9], delta[4], gam_max, gam_min, gridnoise, hourglass[2], hypocenter[3], rcrit,
    rexpand, rho_max, rho_min, slipvector[3], svtol, tm0, trelax, vdamp, vp_max, vp_min, vpml, vrup,
    vs_max, vs_min;
static char source[17], faultnormal[3];
static real dt, dx[3], tm = 0;
static int it = 0, nt, ifn, irup;
static int nn[3], nm[3], nl3[3], nhalo[3], i1bc[3], i2bc[3], i1core[3], i2core[3], i1node[3],
    i2node[3], i1cell[3], i2cell[3], nnoff[3];

typedef struct {
    real *p;
    int n[3];
} A3; /* Fortran-ordered 3-D array view, 1-based indices */

#define AT(a, j, k, l)                                                                    \
    ((a).p[((size_t)(j) - 1) +                                                            \
           (size_t)(a).n[0] * (((size_t)(k) - 1) + (size_t)(a).n[1] * ((size_t)(l) - 1))])

static size_t vol_n;                                /* nm(1)*nm(2)*nm(3) */
static real *dn1, *dn2, *dc1, *dc2;                 /* 1-based: [p-1] */
static real *p1, *p2, *p3, *p4, *p5, *p6, *g1, *g2, *g3, *g4, *g5, *g6;
static int pn[6][3], gn[6][3];                       /* slab array shapes (without comp) */
static A3 vc, mr, lam, mu, gam, yy, s1, s2;
static real *xx, *vv, *uu, *w1, *w2;                 /* 4-D: (nm, 3) */
/* fault */
static int fm[3];
static A3 mus, mud, dc, co, area, rhypo, lamf, muf, sl, psv, trup, tarr, tn, ts, f1, f2;
static real *nhat, *t0, *t1, *t2, *t3;
/* statistics: sord/statistics.f90:3-9 */
static int amaxloc[3], vmaxloc[3], umaxloc[3], wmaxloc[3];
static real amax, vmax, umax, wmax, samax, svmax, sumax, slmax, tsmax, tnmin, tnmax, tarrmax, efric,
    estrain, moment;
/* timers: sord/utilities.f90:3 */
static int64_t timers[4];
static struct timespec clock0;

static int64_t clock_ms(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (int64_t)(t.tv_sec - clock0.tv_sec) * 1000 + (t.tv_nsec - clock0.tv_nsec) / 1000000;
}

static A3 vol3(real *p) {
    A3 a = {p, {nm[0], nm[1], nm[2]}};
    return a;
}
static A3 comp(real *p, int i) { /* component i (1-based) of a 4-D volume array */
    return vol3(p + vol_n * (size_t)(i - 1));
}
static A3 fcomp(real *p, int i) {
    size_t n = (size_t)fm[0] * fm[1] * fm[2];
    A3 a = {p + n * (size_t)(i - 1), {fm[0], fm[1], fm[2]}};
    return a;
}
static real *ralloc(size_t n) {
    real *p = (real *)calloc(n ? n : 1, sizeof(real));
    if (!p) die(2, "out of memory");
    return p;
}
static int imax(int a, int b) { return a > b ? a : b; }
static real fmaxr(real a, real b) { return a > b ? a : b; }
static real fminr(real a, real b) { return a < b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }
/* Fortran modulo for positive divisor */
static int fmodulo(int a, int p) {
    int r = a % p;
    return r < 0 ? r + p : r;
}

/* ------------------------------------------------------------------------ */
/* utilities: sord/utilities.f90                                            */
/* ------------------------------------------------------------------------ */

/* rinvert: sord/utilities.f90:40-50 */
static void rinvert(real *f, size_t n) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++)
        if (f[i] != 0) f[i] = R(1.0) / f[i];
}

/* rlimits: sord/utilities.f90:52-77 */
static void rlimits(real *f, real fmin, real fmax, size_t n) {
    if (fmin > 0 && fmax > 0) {
        for (size_t i = 0; i < n; i++) f[i] = fmaxr(fminr(f[i], fmax), fmin);
    } else if (fmin > 0) {
        for (size_t i = 0; i < n; i++) f[i] = fmaxr(f[i], fmin);
    } else if (fmax > 0) {
        for (size_t i = 0; i < n; i++) f[i] = fminr(f[i], fmax);
    }
}

/* set_halo: sord/utilities.f90:103-118 */
static void set_halo(A3 f, real r, const int i1[3], const int i2[3]) {
    int i3[3], i4[3];
    for (int i = 0; i < 3; i++) {
        i3[i] = imin(i1[i], f.n[i] + 1);
        i4[i] = imax(i2[i], 0);
    }
    for (int l = 1; l <= f.n[2]; l++)
        for (int k = 1; k <= f.n[1]; k++)
            for (int j = 1; j <= f.n[0]; j++) {
                int out = (f.n[0] > 1 && (j < i3[0] || j > i4[0])) ||
                          (f.n[1] > 1 && (k < i3[1] || k > i4[1])) ||
                          (f.n[2] > 1 && (l < i3[2] || l > i4[2]));
                if (out) AT(f, j, k, l) = r;
            }
}

/* average: sord/utilities.f90:79-101 */
static void average(A3 f2_, A3 f1_, const int i1[3], const int i2[3], int d) {
    for (int i = 0; i < 3; i++)
        if (i1[i] < 1 || i2[i] > f1_.n[i]) die(2, "error in average");
#pragma omp parallel for schedule(static)
    for (int l = i1[2]; l <= i2[2]; l++)
        for (int k = i1[1]; k <= i2[1]; k++)
            for (int j = i1[0]; j <= i2[0]; j++)
                AT(f2_, j, k, l) =
                    R(0.125) * (AT(f1_, j, k, l) + AT(f1_, j + d, k + d, l + d) +
                                AT(f1_, j, k + d, l + d) + AT(f1_, j + d, k, l) +
                                AT(f1_, j + d, k, l + d) + AT(f1_, j, k + d, l) +
                                AT(f1_, j + d, k + d, l) + AT(f1_, j, k, l + d));
    set_halo(f2_, 0, i1, i2);
}

/* vector_norm: sord/utilities.f90:120-140 */
static void vector_norm(A3 f, real *w, const int i1[3], const int i2[3], const int di[3]) {
    A3 a = comp(w, 1), b = comp(w, 2), c = comp(w, 3);
#pragma omp parallel for schedule(static)
    for (int l = i1[2]; l <= i2[2]; l += di[2])
        for (int k = i1[1]; k <= i2[1]; k += di[1])
            for (int j = i1[0]; j <= i2[0]; j += di[0])
                AT(f, j, k, l) = AT(a, j, k, l) * AT(a, j, k, l) + AT(b, j, k, l) * AT(b, j, k, l) +
                                 AT(c, j, k, l) * AT(c, j, k, l);
}

/* tensor_norm: sord/utilities.f90:142-166 */
static void tensor_norm(A3 f, real *wa, real *wb, const int i1[3], const int i2[3], const int di[3]) {
    A3 a1 = comp(wa, 1), a2 = comp(wa, 2), a3 = comp(wa, 3);
    A3 b1 = comp(wb, 1), b2 = comp(wb, 2), b3 = comp(wb, 3);
#pragma omp parallel for schedule(static)
    for (int l = i1[2]; l <= i2[2]; l += di[2])
        for (int k = i1[1]; k <= i2[1]; k += di[1])
            for (int j = i1[0]; j <= i2[0]; j += di[0])
                AT(f, j, k, l) = (AT(a1, j, k, l) * AT(a1, j, k, l) + AT(a2, j, k, l) * AT(a2, j, k, l) +
                                  AT(a3, j, k, l) * AT(a3, j, k, l)) +
                                 R(2.0) * (AT(b1, j, k, l) * AT(b1, j, k, l) +
                                           AT(b2, j, k, l) * AT(b2, j, k, l) +
                                           AT(b3, j, k, l) * AT(b3, j, k, l));
}

/* time_function: sord/utilities.f90:223-282 */
static real time_function(const char *pulse, real t, real dt_, real tau) {
    const real pi = R(3.14159265);
    real f = 0, a, b;
    if (!strcmp(pulse, "const")) {
        f = 1;
    } else if (!strcmp(pulse, "delta")) {
        if (FABS(t) < R(0.25) * dt_) f = R(1.0) / dt_;
    } else if (!strcmp(pulse, "step") || !strcmp(pulse, "integral_delta")) {
        if (FABS(t) < R(0.25) * dt_)
            f = R(0.5);
        else if (t >= R(0.25) * dt_)
            f = 1;
    } else if (!strcmp(pulse, "brune")) {
        if (0 < t) {
            a = R(1.0) / tau;
            f = EXP(-a * t) * a * a * t;
        }
    } else if (!strcmp(pulse, "integral_brune")) {
        if (0 < t) {
            a = R(1.0) / tau;
            f = R(1.0) - EXP(-a * t) * (a * t + R(1.0));
        }
    } else if (!strcmp(pulse, "hann")) {
        b = pi * tau;
        if (-b < t && t < b) {
            a = R(1.0) / tau;
            f = R(0.5) / pi * a * (R(1.0) + COS(a * t));
        }
    } else if (!strcmp(pulse, "integral_hann")) {
        b = pi * tau;
        if (-b < t && t < b) {
            a = R(1.0) / tau;
            f = R(0.5) + R(0.5) / pi * (a * t + SIN(a * t));
        } else if (0 < t) {
            f = 1;
        }
    } else if (!strcmp(pulse, "gaussian") || !strcmp(pulse, "integral_ricker1")) {
        a = R(0.5) / (tau * tau);
        b = SQRT(a / pi);
        f = EXP(-a * t * t) * b;
    } else if (!strcmp(pulse, "ricker1") || !strcmp(pulse, "integral_ricker2")) {
        a = R(0.5) / (tau * tau);
        b = SQRT(a / pi) * R(2.0) * a;
        f = -EXP(-a * t * t) * b * t;
    } else if (!strcmp(pulse, "ricker2")) {
        a = R(0.5) / (tau * tau);
        b = SQRT(a / pi) * R(4.0) * a;
        f = EXP(-a * t * t) * b * (a * t * t - R(0.5));
    } else {
        fprintf(stderr, "invalid time func: %s\n", pulse);
        exit(1);
    }
    return f;
}

/* ------------------------------------------------------------------------ */
/* boundary conditions: sord/boundary_cond.f90                              */
/* ------------------------------------------------------------------------ */

static int fmod10(int a) { return a % 10; } /* Fortran mod: sign of dividend, as C */

/* scalar_bc: sord/boundary_cond.f90:7-27 */
static void scalar_bc(A3 f, const int bca[3], const int bcb[3], const int i1[3], const int i2[3]) {
    int b1[3], b2[3];
    for (int i = 0; i < 3; i++) {
        b1[i] = abs(fmod10(bca[i]));
        b2[i] = abs(fmod10(bcb[i]));
        if (b1[i] == 4) b1[i] = 2;
        if (b2[i] == 4) b2[i] = 2;
        if (f.n[i] == 1 || i1[i] < 1 || i1[i] >= f.n[i]) b1[i] = 99;
        if (f.n[i] == 1 || i2[i] <= 1 || i2[i] > f.n[i]) b2[i] = 99;
    }
    int j1 = i1[0], j2 = i2[0], k1 = i1[1], k2 = i2[1], l1 = i1[2], l2 = i2[2];
    if (b1[0] == 2)
        for (int l = 1; l <= f.n[2]; l++)
            for (int k = 1; k <= f.n[1]; k++) AT(f, j1, k, l) = AT(f, j1 + 1, k, l);
    if (b1[1] == 2)
        for (int l = 1; l <= f.n[2]; l++)
            for (int j = 1; j <= f.n[0]; j++) AT(f, j, k1, l) = AT(f, j, k1 + 1, l);
    if (b1[2] == 2)
        for (int k = 1; k <= f.n[1]; k++)
            for (int j = 1; j <= f.n[0]; j++) AT(f, j, k, l1) = AT(f, j, k, l1 + 1);
    if (b2[0] == 2)
        for (int l = 1; l <= f.n[2]; l++)
            for (int k = 1; k <= f.n[1]; k++) AT(f, j2, k, l) = AT(f, j2 - 1, k, l);
    if (b2[1] == 2)
        for (int l = 1; l <= f.n[2]; l++)
            for (int j = 1; j <= f.n[0]; j++) AT(f, j, k2, l) = AT(f, j, k2 - 1, l);
    if (b2[2] == 2)
        for (int k = 1; k <= f.n[1]; k++)
            for (int j = 1; j <= f.n[0]; j++) AT(f, j, k, l2) = AT(f, j, k, l2 - 1);
}

/* set plane `ix` (0-based axis `ax`) of component array f to value */
static void plane_set(A3 f, int ax, int ix, real v) {
    int lo[3] = {1, 1, 1}, hi[3] = {f.n[0], f.n[1], f.n[2]};
    lo[ax] = hi[ax] = ix;
    for (int l = lo[2]; l <= hi[2]; l++)
        for (int k = lo[1]; k <= hi[1]; k++)
            for (int j = lo[0]; j <= hi[0]; j++) AT(f, j, k, l) = v;
}
/* f(plane dst) = s * g(plane src) along axis ax */
static void plane_copy(A3 f, int ax, int dst, A3 g, int src, real s) {
    int lo[3] = {1, 1, 1}, hi[3] = {f.n[0], f.n[1], f.n[2]};
    lo[ax] = hi[ax] = dst;
    for (int l = lo[2]; l <= hi[2]; l++)
        for (int k = lo[1]; k <= hi[1]; k++)
            for (int j = lo[0]; j <= hi[0]; j++) {
                int jj = j, kk = k, ll = l;
                if (ax == 0) jj = src;
                if (ax == 1) kk = src;
                if (ax == 2) ll = src;
                AT(f, j, k, l) = s * AT(g, jj, kk, ll);
            }
}

/* vector_bc: sord/boundary_cond.f90:30-115.  f is a 4-D (n, 3) array. */
static void vector_bc(real *f, const int n[3], const int bca[3], const int bcb[3], const int i1[3],
                      const int i2[3]) {
    size_t N = (size_t)n[0] * n[1] * n[2];
    A3 c[3];
    for (int i = 0; i < 3; i++) {
        c[i].p = f + N * i;
        c[i].n[0] = n[0];
        c[i].n[1] = n[1];
        c[i].n[2] = n[2];
    }
    int b1[3], b2[3], sa[3], sb[3];
    for (int i = 0; i < 3; i++) {
        b1[i] = fmod10(bca[i]);
        b2[i] = fmod10(bcb[i]);
        if (n[i] == 1 || i1[i] < 1 || i1[i] > n[i]) b1[i] = 99;
        if (n[i] == 1 || i2[i] < 1 || i2[i] > n[i]) b2[i] = 99;
    }
    /* anti-mirror symmetry at the node */
    for (int a = 0; a < 3; a++) {
        int b = (a + 1) % 3, cc = (a + 2) % 3;
        if (b1[a] == -1) {
            plane_set(c[b], a, i1[a], 0);
            plane_set(c[cc], a, i1[a], 0);
        }
    }
    for (int a = 0; a < 3; a++) {
        int b = (a + 1) % 3, cc = (a + 2) % 3;
        if (b2[a] == -1) {
            plane_set(c[b], a, i2[a], 0);
            plane_set(c[cc], a, i2[a], 0);
        }
    }
    /* mirror symmetry at the node */
    for (int a = 0; a < 3; a++)
        if (b1[a] == 1) plane_set(c[a], a, i1[a], 0);
    for (int a = 0; a < 3; a++)
        if (b2[a] == 1) plane_set(c[a], a, i2[a], 0);
    /* rigid */
    for (int a = 0; a < 3; a++)
        if (b1[a] == 3)
            for (int i = 0; i < 3; i++) plane_set(c[i], a, i1[a], 0);
    for (int a = 0; a < 3; a++)
        if (b2[a] == 3)
            for (int i = 0; i < 3; i++) plane_set(c[i], a, i2[a], 0);
    for (int i = 0; i < 3; i++) {
        if (i1[i] >= n[i]) b1[i] = 99;
        if (i2[i] <= 1) b2[i] = 99;
    }
    /* continue */
    for (int a = 0; a < 3; a++)
        if (b1[a] == 4)
            for (int i = 0; i < 3; i++) plane_copy(c[i], a, i1[a], c[i], i1[a] + 1, 1);
    for (int a = 0; a < 3; a++)
        if (b2[a] == 4)
            for (int i = 0; i < 3; i++) plane_copy(c[i], a, i2[a], c[i], i2[a] - 1, 1);
    /* symmetry at the cell */
    for (int i = 0; i < 3; i++) {
        b1[i] = abs(b1[i]);
        b2[i] = abs(b2[i]);
        sa[i] = bca[i] >= 0 ? 1 : -1;
        sb[i] = bcb[i] >= 0 ? 1 : -1;
    }
    for (int a = 0; a < 3; a++)
        if (b1[a] == 2)
            for (int i = 0; i < 3; i++)
                plane_copy(c[i], a, i1[a], c[i], i1[a] + 1, (real)(i == a ? -sa[a] : sa[a]));
    for (int a = 0; a < 3; a++)
        if (b2[a] == 2)
            for (int i = 0; i < 3; i++)
                plane_copy(c[i], a, i2[a], c[i], i2[a] - 1, (real)(i == a ? -sb[a] : sb[a]));
}

/* ------------------------------------------------------------------------ */
/* support operators                                                        */
/* ------------------------------------------------------------------------ */

/* diff_nc: node to cell, sord/diff_nc_op.f90:6-73 */
static void diff_nc(A3 df, real *f4, int i, int a, const int i1[3], const int i2[3], real *x4) {
    if (i1[0] > i2[0] || i1[1] > i2[1] || i1[2] > i2[2]) return;
    real pdx = dx[0] * dx[1] * dx[2];
    real h = pdx >= 0 ? R(1.0) / R(12.0) : -(R(1.0) / R(12.0));
    int b = a % 3 + 1, c = (a + 1) % 3 + 1;
    A3 f = comp(f4, i), xb = comp(x4, b), xc = comp(x4, c);
#define F(j, k, l) AT(f, j, k, l)
#define B(j, k, l) AT(xb, j, k, l)
#define C(j, k, l) AT(xc, j, k, l)
#pragma omp parallel for schedule(static)
    for (int l = i1[2]; l <= i2[2]; l++)
        for (int k = i1[1]; k <= i2[1]; k++) {
            for (int j = i1[0]; j <= i2[0]; j++)
                AT(df, j, k, l) = h * F(j, k, l) *
                    ((B(j, k + 1, l + 1) - B(j + 1, k, l)) * (C(j, k + 1, l) - C(j, k, l + 1)) + B(j + 1, k, l) * (C(j + 1, k, l + 1) - C(j + 1, k + 1, l))
                   + (B(j + 1, k, l + 1) - B(j, k + 1, l)) * (C(j, k, l + 1) - C(j + 1, k, l)) + B(j, k + 1, l) * (C(j + 1, k + 1, l) - C(j, k + 1, l + 1))
                   + (B(j + 1, k + 1, l) - B(j, k, l + 1)) * (C(j + 1, k, l) - C(j, k + 1, l)) + B(j, k, l + 1) * (C(j, k + 1, l + 1) - C(j + 1, k, l + 1)));
            for (int j = i1[0]; j <= i2[0]; j++)
                AT(df, j, k, l) = AT(df, j, k, l) + h * F(j + 1, k, l) *
                    ((B(j + 1, k + 1, l + 1) - B(j, k, l)) * (C(j + 1, k, l + 1) - C(j + 1, k + 1, l)) + B(j, k, l) * (C(j, k + 1, l) - C(j, k, l + 1))
                   + (B(j, k + 1, l) - B(j + 1, k, l + 1)) * (C(j + 1, k + 1, l) - C(j, k, l)) + B(j + 1, k, l + 1) * (C(j, k, l + 1) - C(j + 1, k + 1, l + 1))
                   + (B(j, k, l + 1) - B(j + 1, k + 1, l)) * (C(j, k, l) - C(j + 1, k, l + 1)) + B(j + 1, k + 1, l) * (C(j + 1, k + 1, l + 1) - C(j, k + 1, l)));
            for (int j = i1[0]; j <= i2[0]; j++)
                AT(df, j, k, l) = AT(df, j, k, l) + h * F(j, k + 1, l) *
                    ((B(j + 1, k + 1, l + 1) - B(j, k, l)) * (C(j + 1, k + 1, l) - C(j, k + 1, l + 1)) + B(j, k, l) * (C(j, k, l + 1) - C(j + 1, k, l))
                   + (B(j + 1, k, l) - B(j, k + 1, l + 1)) * (C(j, k, l) - C(j + 1, k + 1, l)) + B(j, k + 1, l + 1) * (C(j + 1, k + 1, l + 1) - C(j, k, l + 1))
                   + (B(j, k, l + 1) - B(j + 1, k + 1, l)) * (C(j, k + 1, l + 1) - C(j, k, l)) + B(j + 1, k + 1, l) * (C(j + 1, k, l) - C(j + 1, k + 1, l + 1)));
            for (int j = i1[0]; j <= i2[0]; j++)
                AT(df, j, k, l) = AT(df, j, k, l) + h * F(j + 1, k + 1, l) *
                    ((B(j, k, l) - B(j + 1, k + 1, l + 1)) * (C(j, k + 1, l) - C(j + 1, k, l)) + B(j + 1, k + 1, l + 1) * (C(j + 1, k, l + 1) - C(j, k + 1, l + 1))
                   + (B(j, k + 1, l + 1) - B(j + 1, k, l)) * (C(j + 1, k + 1, l + 1) - C(j, k + 1, l)) + B(j + 1, k, l) * (C(j, k, l) - C(j + 1, k, l + 1))
                   + (B(j + 1, k, l + 1) - B(j, k + 1, l)) * (C(j + 1, k, l) - C(j + 1, k + 1, l + 1)) + B(j, k + 1, l) * (C(j, k + 1, l + 1) - C(j, k, l)));
            for (int j = i1[0]; j <= i2[0]; j++)
                AT(df, j, k, l) = AT(df, j, k, l) + h * F(j, k, l + 1) *
                    ((B(j + 1, k + 1, l + 1) - B(j, k, l)) * (C(j, k + 1, l + 1) - C(j + 1, k, l + 1)) + B(j, k, l) * (C(j + 1, k, l) - C(j, k + 1, l))
                   + (B(j + 1, k, l) - B(j, k + 1, l + 1)) * (C(j + 1, k, l + 1) - C(j, k, l)) + B(j, k + 1, l + 1) * (C(j, k + 1, l) - C(j + 1, k + 1, l + 1))
                   + (B(j, k + 1, l) - B(j + 1, k, l + 1)) * (C(j, k, l) - C(j, k + 1, l + 1)) + B(j + 1, k, l + 1) * (C(j + 1, k + 1, l + 1) - C(j + 1, k, l)));
            for (int j = i1[0]; j <= i2[0]; j++)
                AT(df, j, k, l) = AT(df, j, k, l) + h * F(j + 1, k, l + 1) *
                    ((B(j, k, l) - B(j + 1, k + 1, l + 1)) * (C(j + 1, k, l) - C(j, k, l + 1)) + B(j + 1, k + 1, l + 1) * (C(j, k + 1, l + 1) - C(j + 1, k + 1, l))
                   + (B(j, k + 1, l + 1) - B(j + 1, k, l)) * (C(j, k, l + 1) - C(j + 1, k + 1, l + 1)) + B(j + 1, k, l) * (C(j + 1, k + 1, l) - C(j, k, l))
                   + (B(j + 1, k + 1, l) - B(j, k, l + 1)) * (C(j + 1, k + 1, l + 1) - C(j + 1, k, l)) + B(j, k, l + 1) * (C(j, k, l) - C(j, k + 1, l + 1)));
            for (int j = i1[0]; j <= i2[0]; j++)
                AT(df, j, k, l) = AT(df, j, k, l) + h * F(j + 1, k + 1, l + 1) *
                    ((B(j + 1, k, l) - B(j, k + 1, l + 1)) * (C(j + 1, k + 1, l) - C(j + 1, k, l + 1)) + B(j, k + 1, l + 1) * (C(j, k, l + 1) - C(j, k + 1, l))
                   + (B(j, k + 1, l) - B(j + 1, k, l + 1)) * (C(j, k + 1, l + 1) - C(j + 1, k + 1, l)) + B(j + 1, k, l + 1) * (C(j + 1, k, l) - C(j, k, l + 1))
                   + (B(j, k, l + 1) - B(j + 1, k + 1, l)) * (C(j + 1, k, l + 1) - C(j, k + 1, l + 1)) + B(j + 1, k + 1, l) * (C(j, k + 1, l) - C(j + 1, k, l)));
            for (int j = i1[0]; j <= i2[0]; j++)
                AT(df, j, k, l) = AT(df, j, k, l) + h * F(j, k + 1, l + 1) *
                    ((B(j, k, l) - B(j + 1, k + 1, l + 1)) * (C(j, k, l + 1) - C(j, k + 1, l)) + B(j + 1, k + 1, l + 1) * (C(j + 1, k + 1, l) - C(j + 1, k, l + 1))
                   + (B(j + 1, k, l + 1) - B(j, k + 1, l)) * (C(j + 1, k + 1, l + 1) - C(j, k, l + 1)) + B(j, k + 1, l) * (C(j, k, l) - C(j + 1, k + 1, l))
                   + (B(j + 1, k + 1, l) - B(j, k, l + 1)) * (C(j, k + 1, l) - C(j + 1, k + 1, l + 1)) + B(j, k, l + 1) * (C(j + 1, k, l + 1) - C(j, k, l)));
        }
}

/* diff_cn: cell to node, sord/diff_cn_op.f90:6-73 */
static void diff_cn(A3 df, real *f4, int i, int a, const int i1[3], const int i2[3], real *x4) {
    if (i1[0] > i2[0] || i1[1] > i2[1] || i1[2] > i2[2]) return;
    real pdx = dx[0] * dx[1] * dx[2];
    real h = pdx >= 0 ? R(1.0) / R(12.0) : -(R(1.0) / R(12.0));
    int b = a % 3 + 1, c = (a + 1) % 3 + 1;
    A3 f = comp(f4, i), xb = comp(x4, b), xc = comp(x4, c);
#pragma omp parallel for schedule(static)
    for (int l = i1[2]; l <= i2[2]; l++)
        for (int k = i1[1]; k <= i2[1]; k++) {
            for (int j = i1[0]; j <= i2[0]; j++)
                AT(df, j, k, l) = h * F(j - 1, k - 1, l - 1) *
                    ((B(j - 1, k, l) - B(j, k - 1, l - 1)) * (C(j, k, l - 1) - C(j, k - 1, l)) + B(j - 1, k, l) * (C(j - 1, k, l - 1) - C(j - 1, k - 1, l))
                   + (B(j, k - 1, l) - B(j - 1, k, l - 1)) * (C(j - 1, k, l) - C(j, k, l - 1)) + B(j, k - 1, l) * (C(j - 1, k - 1, l) - C(j, k - 1, l - 1))
                   + (B(j, k, l - 1) - B(j - 1, k - 1, l)) * (C(j, k - 1, l) - C(j - 1, k, l)) + B(j, k, l - 1) * (C(j, k - 1, l - 1) - C(j - 1, k, l - 1)));
            for (int j = i1[0]; j <= i2[0]; j++)
                AT(df, j, k, l) = AT(df, j, k, l) + h * F(j, k - 1, l - 1) *
                    ((B(j + 1, k, l) - B(j, k - 1, l - 1)) * (C(j, k - 1, l) - C(j, k, l - 1)) + B(j + 1, k, l) * (C(j + 1, k - 1, l) - C(j + 1, k, l - 1))
                   + (B(j, k - 1, l) - B(j + 1, k, l - 1)) * (C(j, k, l - 1) - C(j + 1, k, l)) + B(j, k - 1, l) * (C(j, k - 1, l - 1) - C(j + 1, k - 1, l))
                   + (B(j, k, l - 1) - B(j + 1, k - 1, l)) * (C(j + 1, k, l) - C(j, k - 1, l)) + B(j, k, l - 1) * (C(j + 1, k, l - 1) - C(j, k - 1, l - 1)));
            for (int j = i1[0]; j <= i2[0]; j++)
                AT(df, j, k, l) = AT(df, j, k, l) + h * F(j - 1, k - 1, l) *
                    ((B(j, k, l + 1) - B(j - 1, k - 1, l)) * (C(j - 1, k, l) - C(j, k - 1, l)) + B(j, k, l + 1) * (C(j - 1, k, l + 1) - C(j, k - 1, l + 1))
                   + (B(j - 1, k, l) - B(j, k - 1, l + 1)) * (C(j, k - 1, l) - C(j, k, l + 1)) + B(j - 1, k, l) * (C(j - 1, k - 1, l) - C(j - 1, k, l + 1))
                   + (B(j, k - 1, l) - B(j - 1, k, l + 1)) * (C(j, k, l + 1) - C(j - 1, k, l)) + B(j, k - 1, l) * (C(j, k - 1, l + 1) - C(j - 1, k - 1, l)));
            for (int j = i1[0]; j <= i2[0]; j++)
                AT(df, j, k, l) = AT(df, j, k, l) + h * F(j, k - 1, l) *
                    ((B(j, k - 1, l) - B(j + 1, k, l + 1)) * (C(j + 1, k, l) - C(j, k, l + 1)) + B(j, k - 1, l) * (C(j + 1, k - 1, l) - C(j, k - 1, l + 1))
                   + (B(j, k, l + 1) - B(j + 1, k - 1, l)) * (C(j, k - 1, l) - C(j + 1, k, l)) + B(j, k, l + 1) * (C(j, k - 1, l + 1) - C(j + 1, k, l + 1))
                   + (B(j + 1, k, l) - B(j, k - 1, l + 1)) * (C(j, k, l + 1) - C(j, k - 1, l)) + B(j + 1, k, l) * (C(j + 1, k, l + 1) - C(j + 1, k - 1, l)));
            for (int j = i1[0]; j <= i2[0]; j++)
                AT(df, j, k, l) = AT(df, j, k, l) + h * F(j - 1, k, l - 1) *
                    ((B(j, k + 1, l) - B(j - 1, k, l - 1)) * (C(j, k, l - 1) - C(j - 1, k, l)) + B(j, k + 1, l) * (C(j, k + 1, l - 1) - C(j - 1, k + 1, l))
                   + (B(j, k, l - 1) - B(j - 1, k + 1, l)) * (C(j - 1, k, l) - C(j, k + 1, l)) + B(j, k, l - 1) * (C(j - 1, k, l - 1) - C(j, k + 1, l - 1))
                   + (B(j - 1, k, l) - B(j, k + 1, l - 1)) * (C(j, k + 1, l) - C(j, k, l - 1)) + B(j - 1, k, l) * (C(j - 1, k + 1, l) - C(j - 1, k, l - 1)));
            for (int j = i1[0]; j <= i2[0]; j++)
                AT(df, j, k, l) = AT(df, j, k, l) + h * F(j, k, l - 1) *
                    ((B(j, k, l - 1) - B(j + 1, k + 1, l)) * (C(j, k + 1, l) - C(j + 1, k, l)) + B(j, k, l - 1) * (C(j, k + 1, l - 1) - C(j + 1, k, l - 1))
                   + (B(j + 1, k, l) - B(j, k + 1, l - 1)) * (C(j, k, l - 1) - C(j, k + 1, l)) + B(j + 1, k, l) * (C(j + 1, k, l - 1) - C(j + 1, k + 1, l))
                   + (B(j, k + 1, l) - B(j + 1, k, l - 1)) * (C(j + 1, k, l) - C(j, k, l - 1)) + B(j, k + 1, l) * (C(j + 1, k + 1, l) - C(j, k + 1, l - 1)));
            for (int j = i1[0]; j <= i2[0]; j++)
                AT(df, j, k, l) = AT(df, j, k, l) + h * F(j - 1, k, l) *
                    ((B(j - 1, k, l) - B(j, k + 1, l + 1)) * (C(j, k, l + 1) - C(j, k + 1, l)) + B(j - 1, k, l) * (C(j - 1, k, l + 1) - C(j - 1, k + 1, l))
                   + (B(j, k + 1, l) - B(j - 1, k, l + 1)) * (C(j - 1, k, l) - C(j, k, l + 1)) + B(j, k + 1, l) * (C(j - 1, k + 1, l) - C(j, k + 1, l + 1))
                   + (B(j, k, l + 1) - B(j - 1, k + 1, l)) * (C(j, k + 1, l) - C(j - 1, k, l)) + B(j, k, l + 1) * (C(j, k + 1, l + 1) - C(j - 1, k, l + 1)));
            for (int j = i1[0]; j <= i2[0]; j++)
                AT(df, j, k, l) = AT(df, j, k, l) + h * F(j, k, l) *
                    ((B(j + 1, k, l) - B(j, k + 1, l + 1)) * (C(j, k + 1, l) - C(j, k, l + 1)) + B(j + 1, k, l) * (C(j + 1, k + 1, l) - C(j + 1, k, l + 1))
                   + (B(j, k + 1, l) - B(j + 1, k, l + 1)) * (C(j, k, l + 1) - C(j + 1, k, l)) + B(j, k + 1, l) * (C(j, k + 1, l + 1) - C(j + 1, k + 1, l))
                   + (B(j, k, l + 1) - B(j + 1, k + 1, l)) * (C(j + 1, k, l) - C(j, k + 1, l)) + B(j, k, l + 1) * (C(j + 1, k, l + 1) - C(j, k + 1, l + 1)));
        }
#undef F
#undef B
#undef C
}

/* hourglass_nc: node to cell, sord/hourglass_op.f90:7-74.
 * Corner signs of the four modes for corners (dj,dk,dl) in {0,1}^3. */
static const signed char hg_sign[4][8] = {
    /* order: (0,0,0) (1,0,0) (0,1,0) (1,1,0) (0,0,1) (1,0,1) (0,1,1) (1,1,1) */
    {+1, +1, -1, -1, -1, -1, +1, +1},
    {+1, -1, +1, -1, -1, +1, -1, +1},
    {+1, -1, -1, +1, +1, -1, -1, +1},
    {+1, -1, -1, +1, -1, +1, +1, -1},
};
static void hourglass_nc(A3 df, real *f4, int iq, int i, const int i1[3], const int i2[3]) {
    if (i1[0] > i2[0] || i1[1] > i2[1] || i1[2] > i2[2]) return;
    A3 f = comp(f4, i);
    const signed char *s = hg_sign[iq - 1];
    /* summation order of the reference: corners (0,0,0) (1,1,1) (0,1,1) (1,0,0)
       (1,0,1) (0,1,0) (1,1,0) (0,0,1) */
    static const int ord[8] = {0, 7, 6, 1, 5, 2, 3, 4};
#pragma omp parallel for schedule(static)
    for (int l = i1[2]; l <= i2[2]; l++)
        for (int k = i1[1]; k <= i2[1]; k++)
            for (int j = i1[0]; j <= i2[0]; j++) {
                real acc = 0;
                for (int q = 0; q < 8; q++) {
                    int c = ord[q], dj = c & 1, dk = (c >> 1) & 1, dl = (c >> 2) & 1;
                    real v = AT(f, j + dj, k + dk, l + dl);
                    acc = q == 0 ? (s[c] > 0 ? v : -v) : (s[c] > 0 ? acc + v : acc - v);
                }
                AT(df, j, k, l) = acc;
            }
}

/* hourglass_cn: cell to node, sord/hourglass_op.f90:79-146 */
static void hourglass_cn(A3 df, A3 f, int iq, const int i1[3], const int i2[3]) {
    if (i1[0] > i2[0] || i1[1] > i2[1] || i1[2] > i2[2]) return;
    const signed char *s = hg_sign[iq - 1];
    static const int ord[8] = {0, 7, 6, 1, 5, 2, 3, 4};
#pragma omp parallel for schedule(static)
    for (int l = i1[2]; l <= i2[2]; l++)
        for (int k = i1[1]; k <= i2[1]; k++)
            for (int j = i1[0]; j <= i2[0]; j++) {
                real acc = 0;
                for (int q = 0; q < 8; q++) {
                    int c = ord[q], dj = c & 1, dk = (c >> 1) & 1, dl = (c >> 2) & 1;
                    real v = AT(f, j - dj, k - dk, l - dl);
                    acc = q == 0 ? (s[c] > 0 ? v : -v) : (s[c] > 0 ? acc + v : acc - v);
                }
                AT(df, j, k, l) = acc;
            }
}

/* node_normals: sord/surf_normals.f90:53-128 (x is a 4-D volume array) */
static void node_normals(real *nh4, real *x4, const int i1[3], const int i2[3], int ihat) {
    size_t nf = (size_t)fm[0] * fm[1] * fm[2];
    for (size_t q = 0; q < 3 * nf; q++) nh4[q] = 0;
    real pdx = dx[0] * dx[1] * dx[2];
    real h = pdx >= 0 ? R(1.0) / R(12.0) : -(R(1.0) / R(12.0));
    for (int a = 1; a <= 3; a++) {
        int b = a % 3 + 1, c = (a + 1) % 3 + 1;
        A3 n = fcomp(nh4, a), xb = comp(x4, b), xc = comp(x4, c);
#define B(j, k, l) AT(xb, j, k, l)
#define C(j, k, l) AT(xc, j, k, l)
        if (ihat == 1) {
            int j = i1[0];
            for (int l = i1[2]; l <= i2[2]; l++)
                for (int k = i1[1]; k <= i2[1]; k++)
                    AT(n, 1, k, l) = h *
                        ( B(j, k + 1, l) * ( C(j, k, l + 1) + C(j, k + 1, l + 1)
                                           - C(j, k, l - 1) - C(j, k + 1, l - 1) )
                        + B(j, k - 1, l) * ( C(j, k, l - 1) + C(j, k - 1, l - 1)
                                           - C(j, k, l + 1) - C(j, k - 1, l + 1) )
                        + B(j, k, l + 1) * ( C(j, k - 1, l) + C(j, k - 1, l + 1)
                                           - C(j, k + 1, l) - C(j, k + 1, l + 1) )
                        + B(j, k, l - 1) * ( C(j, k + 1, l) + C(j, k + 1, l - 1)
                                           - C(j, k - 1, l) - C(j, k - 1, l - 1) )
                        + B(j, k + 1, l + 1) * (C(j, k, l + 1) - C(j, k + 1, l))
                        + B(j, k - 1, l - 1) * (C(j, k, l - 1) - C(j, k - 1, l))
                        + B(j, k - 1, l + 1) * (C(j, k - 1, l) - C(j, k, l + 1))
                        + B(j, k + 1, l - 1) * (C(j, k + 1, l) - C(j, k, l - 1)) );
        } else if (ihat == 2) {
            int k = i1[1];
            for (int l = i1[2]; l <= i2[2]; l++)
                for (int j = i1[0]; j <= i2[0]; j++)
                    AT(n, j, 1, l) = h *
                        ( B(j, k, l + 1) * ( C(j + 1, k, l) + C(j + 1, k, l + 1)
                                           - C(j - 1, k, l) - C(j - 1, k, l + 1) )
                        + B(j, k, l - 1) * ( C(j - 1, k, l) + C(j - 1, k, l - 1)
                                           - C(j + 1, k, l) - C(j + 1, k, l - 1) )
                        + B(j + 1, k, l) * ( C(j, k, l - 1) + C(j + 1, k, l - 1)
                                           - C(j, k, l + 1) - C(j + 1, k, l + 1) )
                        + B(j - 1, k, l) * ( C(j, k, l + 1) + C(j - 1, k, l + 1)
                                           - C(j, k, l - 1) - C(j - 1, k, l - 1) )
                        + B(j + 1, k, l + 1) * (C(j + 1, k, l) - C(j, k, l + 1))
                        + B(j - 1, k, l - 1) * (C(j - 1, k, l) - C(j, k, l - 1))
                        + B(j + 1, k, l - 1) * (C(j, k, l - 1) - C(j + 1, k, l))
                        + B(j - 1, k, l + 1) * (C(j, k, l + 1) - C(j - 1, k, l)) );
        } else if (ihat == 3) {
            int l = i1[2];
            for (int k = i1[1]; k <= i2[1]; k++)
                for (int j = i1[0]; j <= i2[0]; j++)
                    AT(n, j, k, 1) = h *
                        ( B(j + 1, k, l) * ( C(j, k + 1, l) + C(j + 1, k + 1, l)
                                           - C(j, k - 1, l) - C(j + 1, k - 1, l) )
                        + B(j - 1, k, l) * ( C(j, k - 1, l) + C(j - 1, k - 1, l)
                                           - C(j, k + 1, l) - C(j - 1, k + 1, l) )
                        + B(j, k + 1, l) * ( C(j - 1, k, l) + C(j - 1, k + 1, l)
                                           - C(j + 1, k, l) - C(j + 1, k + 1, l) )
                        + B(j, k - 1, l) * ( C(j + 1, k, l) + C(j + 1, k - 1, l)
                                           - C(j - 1, k, l) - C(j - 1, k - 1, l) )
                        + B(j + 1, k + 1, l) * (C(j, k + 1, l) - C(j + 1, k, l))
                        + B(j - 1, k - 1, l) * (C(j, k - 1, l) - C(j - 1, k, l))
                        + B(j - 1, k + 1, l) * (C(j - 1, k, l) - C(j, k + 1, l))
                        + B(j + 1, k - 1, l) * (C(j + 1, k, l) - C(j, k - 1, l)) );
        } else {
            die(2, "error: surfnormal");
        }
#undef B
#undef C
    }
}

/* ------------------------------------------------------------------------ */
/* sord.in: sord/input_output.f90:21-98                                     */
/* ------------------------------------------------------------------------ */
typedef struct {
    real x1[3], x2[3], val, tau;
    int ii[4][3]; /* ii[dim][start,stop,step], 1-based starts */
    int nb, ib;
    int fh_open;  /* file written/read so far (records) */
    real *buff;
    char fname[33], field[5], op[3], reg;
} t_io;
static t_io *io_list;
static int nio;

/* next blank- or comma-separated token of a line (list-directed input) */
static char *tok_next(char **s) {
    char *p = *s;
    while (*p == ' ' || *p == '\t' || *p == ',' || *p == '\r') p++;
    if (*p == '\0') {
        *s = p;
        return NULL;
    }
    char *t = p;
    while (*p && *p != ' ' && *p != '\t' && *p != ',' && *p != '\r') p++;
    if (*p) *p++ = '\0';
    *s = p;
    return t;
}

static char *line_tok;
static char *need(void) {
    char *t = tok_next(&line_tok);
    if (!t) die(2, "sord.in: too few values");
    return t;
}
static int rint_(void) { return (int)strtol(need(), NULL, 10); }
static real rreal_(void) {
#if REALSIZE == 8
    return strtod(need(), NULL);
#else
    return strtof(need(), NULL);
#endif
}

static void read_parameters(void) {
    FILE *fp = fopen("sord.in", "rb");
    if (!fp) die(2, "cannot open sord.in");
    fseek(fp, 0, SEEK_END);
    long n = ftell(fp);
    fseek(fp, 0, SEEK_SET);
    char *str = (char *)malloc(n + 2);
    if (fread(str, 1, n, fp) != (size_t)n) die(2, "read error sord.in");
    str[n] = '\0';
    fclose(fp);
    /* split into lines */
    char **lines = (char **)malloc(sizeof(char *) * (n + 2));
    int nlines = 0;
    char *p = str;
    while (*p) {
        lines[nlines++] = p;
        char *e = strchr(p, '\n');
        if (!e) break;
        *e = '\0';
        p = e + 1;
    }
    if (nlines < 1) die(2, "sord.in: empty");
    line_tok = lines[0];
    int nfieldio;
    for (int i = 0; i < 9; i++) affine[i] = rreal_();
    for (int i = 0; i < 3; i++) bc1[i] = rint_();
    for (int i = 0; i < 3; i++) bc2[i] = rint_();
    debug_ = rint_();
    for (int i = 0; i < 4; i++) delta[i] = rreal_();
    {
        char *t = need();
        faultnormal[0] = t[0];
        faultnormal[1] = t[1] ? t[1] : ' ';
        faultnormal[2] = 0;
    }
    faultopening = rint_();
    gam_max = rreal_();
    gam_min = rreal_();
    gridnoise = rreal_();
    for (int i = 0; i < 2; i++) hourglass[i] = rreal_();
    for (int i = 0; i < 3; i++) hypocenter[i] = rreal_();
    for (int i = 0; i < 3; i++) i1pml[i] = rint_();
    for (int i = 0; i < 3; i++) i2pml[i] = rint_();
    itio = rint_();
    itstats = rint_();
    mpin = rint_();
    mpout = rint_();
    for (int i = 0; i < 3; i++) n1expand[i] = rint_();
    for (int i = 0; i < 3; i++) n2expand[i] = rint_();
    nfieldio = rint_();
    npml = rint_();
    for (int i = 0; i < 3; i++) nproc3[i] = rint_();
    nsource = rint_();
    nthread = rint_();
    ppml = rint_();
    rcrit = rreal_();
    rexpand = rreal_();
    rho_max = rreal_();
    rho_min = rreal_();
    for (int i = 0; i < 4; i++) shape_[i] = rint_();
    for (int i = 0; i < 3; i++) slipvector[i] = rreal_();
    {
        char *t = need();
        snprintf(source, sizeof source, "%.16s", t);
    }
    svtol = rreal_();
    tm0 = rreal_();
    trelax = rreal_();
    vdamp = rreal_();
    vp_max = rreal_();
    vp_min = rreal_();
    vpml = rreal_();
    vrup = rreal_();
    vs_max = rreal_();
    vs_min = rreal_();
    /* convert indices: sord/input_output.f90:62 */
    for (int i = 0; i < 3; i++) i1pml[i] += 1;
    if (nfieldio > nlines - 1) die(2, "sord.in: too few field i/o records");
    nio = nfieldio;
    io_list = (t_io *)calloc(nio > 0 ? nio : 1, sizeof(t_io));
    for (int q = 0; q < nio; q++) {
        t_io *io = &io_list[q];
        line_tok = lines[q + 1];
        io->ib = -1;
        snprintf(io->field, sizeof io->field, "%.4s", need());
        io->reg = need()[0];
        for (int d = 0; d < 4; d++)
            for (int e = 0; e < 3; e++) io->ii[d][e] = rint_();
        io->nb = rint_();
        for (int i = 0; i < 3; i++) io->x1[i] = rreal_();
        for (int i = 0; i < 3; i++) io->x2[i] = rreal_();
        io->val = rreal_();
        io->tau = rreal_();
        {
            char *t = need();
            io->op[0] = t[0];
            io->op[1] = t[1] ? t[1] : ' ';
            io->op[2] = 0;
        }
        snprintf(io->fname, sizeof io->fname, "%.32s", need());
        for (int d = 0; d < 4; d++) io->ii[d][0] += 1;
    }
    free(lines);
    free(str);
}

/* ------------------------------------------------------------------------ */
/* setup_dimensions: sord/setup.f90:6-119 (serial process module)           */
/* ------------------------------------------------------------------------ */
static void setup_dimensions(void) {
    for (int i = 0; i < 3; i++) dx[i] = delta[i];
    dt = delta[3];
    for (int i = 0; i < 3; i++) nn[i] = shape_[i];
    nt = shape_[3] > 0 ? shape_[3] : 0;
    switch (faultnormal[1]) {
    case 'o': ifn = 0; break;
    case 'x': ifn = 1; break;
    case 'y': ifn = 2; break;
    case 'z': ifn = 3; break;
    default: die(1, "Error in faultnormal");
    }
    for (int i = 0; i < 3; i++) {
        if (nproc3[i] != 1) die(3, "oracle: serial process module only (nproc3 = 1)");
        nl3[i] = (nn[i] - 1) / nproc3[i] + 1;
        nhalo[i] = 1;
    }
    if (ifn) nhalo[ifn - 1] = 2;
    for (int i = 0; i < 3; i++) {
        nl3[i] = imax(nl3[i], nhalo[i]);
        nnoff[i] = -nhalo[i];
        int nl = imin(nl3[i], nn[i] - nnoff[i] - nhalo[i]);
        nm[i] = nl + 2 * nhalo[i];
        i1bc[i] = 1 - nnoff[i];
        i2bc[i] = nn[i] - nnoff[i];
        i1core[i] = 1 + nhalo[i];
        i2core[i] = nm[i] - nhalo[i];
        i1node[i] = imax(i1bc[i], 2);
        i2node[i] = imin(i2bc[i], nm[i] - 1);
        i1cell[i] = imax(i1bc[i], 1);
        i2cell[i] = imin(i2bc[i] - 1, nm[i] - 1);
        i1pml[i] -= nnoff[i];
        i2pml[i] -= nnoff[i];
    }
    irup = 0;
    if (ifn) {
        irup = (int)floor((double)(hypocenter[ifn - 1] + R(0.000001))) - nnoff[ifn - 1] + 1;
        if (irup + 1 < i1core[ifn - 1] || irup > i2core[ifn - 1]) ifn = 0;
    }
}

/* ------------------------------------------------------------------------ */
/* field i/o: sord/input_output.f90:103-542                                 */
/* ------------------------------------------------------------------------ */
/* sord/utilities.f90:168-221: linear interpolation between nodes read at
 * strides di (the k pass steps the l planes by di(1), as the reference does) */
static void interpolate(A3 f, const int i3[3], const int i4[3], const int di[3]) {
    int i1[3], i2[3];
    for (int a = 0; a < 3; a++) {
        i1[a] = i3[a], i2[a] = i4[a];
        if (i1[a] < 1) i1[a] = i1[a] + (-i1[a] / di[a] + 1) * di[a];
        if (i2[a] > f.n[a]) i2[a] = i1[a] + (f.n[a] - i1[a]) / di[a] * di[a];
    }
    int d = di[0];
    for (int i = 1; i <= d - 1; i++) {
        real h1 = R(1.0) / (real)d * (real)i, h2 = R(1.0) / (real)d * (real)(d - i);
        for (int l = i1[2]; l <= i2[2]; l += di[2])
            for (int k = i1[1]; k <= i2[1]; k += di[1])
                for (int j = i1[0]; j <= i2[0] - d; j += d) AT(f, j + i, k, l) = h1 * AT(f, j, k, l) + h2 * AT(f, j + d, k, l);
    }
    d = di[1];
    for (int i = 1; i <= d - 1; i++) {
        real h1 = R(1.0) / (real)d * (real)i, h2 = R(1.0) / (real)d * (real)(d - i);
        for (int l = i1[2]; l <= i2[2]; l += di[0])
            for (int k = i1[1]; k <= i2[1] - d; k += d)
                for (int j = i1[0]; j <= i2[0]; j++) AT(f, j, k + i, l) = h1 * AT(f, j, k, l) + h2 * AT(f, j, k + d, l);
    }
    d = di[2];
    for (int i = 1; i <= d - 1; i++) {
        real h1 = R(1.0) / (real)d * (real)i, h2 = R(1.0) / (real)d * (real)(d - i);
        for (int l = i1[2]; l <= i2[2] - d; l += d)
            for (int k = i1[1]; k <= i2[1]; k++)
                for (int j = i1[0]; j <= i2[0]; j++) AT(f, j, k, l + i) = h1 * AT(f, j, k, l) + h2 * AT(f, j, k, l + d);
    }
}

static void rio_write(const char *fname, const real *buf, size_t reclen, int o, int n) {
    FILE *fp = fopen(fname, o > 0 ? "r+b" : "wb");
    if (!fp) die(2, "cannot open output file");
    fseek(fp, (long)((size_t)o * reclen * sizeof(real)), SEEK_SET);
    fwrite(buf, sizeof(real), reclen * (size_t)n, fp);
    fclose(fp);
}
static void rio_read(const char *fname, real *buf, size_t reclen, int o, int n) {
    FILE *fp = fopen(fname, "rb");
    if (!fp) {
        fprintf(stderr, "cannot open input file %s\n", fname);
        exit(2);
    }
    fseek(fp, (long)((size_t)o * reclen * sizeof(real)), SEEK_SET);
    if (fread(buf, sizeof(real), reclen * (size_t)n, fp) != reclen * (size_t)n) die(2, "short read");
    fclose(fp);
}

static int fdiv(int a, int b) { return a / b; } /* Fortran integer division (truncation) */

static void field_io(const char *passes, const char *field, A3 f) {
    int64_t tclk = clock_ms();
    timers[1] -= tclk;
    for (const char *pp = passes; *pp; pp++) {
        char pass = *pp;
        for (int q = 0; q < nio; q++) {
            t_io *io = &io_list[q];
            int it1 = io->ii[3][0], it2 = io->ii[3][1], dit = io->ii[3][2];
            int i1[3], i2[3], i3[3], i4[3], di[3], m[4], n[4], o[4];
            real xi[3];
            if (it > it2) continue;
            if (io->op[0] == '#') continue;
            if (it < it1) continue;
            if (fmodulo(it - it1, dit) != 0) continue;
            if (strcmp(field, io->field)) continue;
            if (pass == '<' && io->op[1] == '>') continue;
            if (pass == '>' && io->op[1] != '>') continue;
            for (int i = 0; i < 3; i++) {
                if (io->reg == 'n') {
                    i3[i] = i1node[i];
                    i4[i] = i2node[i];
                    xi[i] = io->x1[i] - (real)nnoff[i] + R(1.0);
                } else {
                    i3[i] = i1cell[i];
                    i4[i] = i2cell[i];
                    xi[i] = io->x1[i] - (real)nnoff[i] + R(0.5);
                }
            }
            if (io->op[0] == '.') {
                for (int i = 0; i < 3; i++) {
                    int fl = (int)floor((double)xi[i]);
                    i1[i] = imax(i3[i], fl);
                    i2[i] = imin(i4[i], fl + 1);
                    di[i] = 1;
                }
            } else {
                for (int i = 0; i < 3; i++) {
                    i1[i] = io->ii[i][0] - nnoff[i];
                    i2[i] = io->ii[i][1] - nnoff[i];
                    di[i] = io->ii[i][2];
                    if (i1[i] < i3[i]) i1[i] = i1[i] + (fdiv(i3[i] - i1[i] - 1, di[i]) + 1) * di[i];
                    if (i2[i] > i4[i]) i2[i] = i1[i] + fdiv(i4[i] - i1[i], di[i]) * di[i];
                }
            }

            /* non-disk input */
            if (io->op[1] != '<' && io->op[1] != '>') {
                if (i2[0] < i1[0] || i2[1] < i1[1] || i2[2] < i1[2]) {
                    io->op[0] = '#';
                    continue;
                }
                real val = io->val * time_function(io->fname, tm, dt, io->tau);
                /* Reference defect D14: for fault fields the reference indexes the
                   size-1 normal axis with the volume index (an out-of-bounds
                   write, sord/input_output.f90:192-199).  The intended plane is
                   used here. */
                if (io->reg == 'f')
                    for (int i = 0; i < 3; i++)
                        if (f.n[i] == 1) i1[i] = i2[i] = 1;
                if (!strcmp(io->op, ". ")) {
                    for (int l = i1[2]; l <= i2[2]; l++)
                        for (int k = i1[1]; k <= i2[1]; k++)
                            for (int j = i1[0]; j <= i2[0]; j++) {
                                real r = (R(1.0) - FABS(xi[0] - (real)j)) *
                                         (R(1.0) - FABS(xi[1] - (real)k)) *
                                         (R(1.0) - FABS(xi[2] - (real)l));
                                AT(f, j, k, l) = AT(f, j, k, l) + val * r;
                            }
                } else if (!strcmp(io->op, "= ") || !strcmp(io->op, "+ ") || !strcmp(io->op, "* ")) {
                    for (int l = i1[2]; l <= i2[2]; l += di[2])
                        for (int k = i1[1]; k <= i2[1]; k += di[1])
                            for (int j = i1[0]; j <= i2[0]; j += di[0]) {
                                if (io->op[0] == '=') AT(f, j, k, l) = val;
                                else if (io->op[0] == '+') AT(f, j, k, l) = AT(f, j, k, l) + val;
                                else AT(f, j, k, l) = AT(f, j, k, l) * val;
                            }
                } else if (io->op[1] == '@' && io->reg != 'f') {
                    /* cube ops: the op on the nodes (cells) of the range whose
                       coordinates (centres) lie in the box [x1, x2].  The
                       reference's set_cube (sord/input_output.f90:546-591,
                       called at :246-251) tests w1 for node fields, which holds
                       the cell centres at initialisation, and w2 for cell
                       fields, which holds no coordinates (defect D9); the
                       intended coordinates are used here, the cell centres by
                       the reference's own 8-node average (average, :79-101). */
                    A3 x[3] = {comp(xx, 1), comp(xx, 2), comp(xx, 3)};
                    for (int l = i1[2]; l <= i2[2]; l += di[2])
                        for (int k = i1[1]; k <= i2[1]; k += di[1])
                            for (int j = i1[0]; j <= i2[0]; j += di[0]) {
                                int in = 1;
                                for (int a = 0; a < 3; a++) {
                                    real c = io->reg == 'n' ? AT(x[a], j, k, l)
                                        : R(0.125) * (AT(x[a], j, k, l) + AT(x[a], j + 1, k + 1, l + 1) +
                                                      AT(x[a], j, k + 1, l + 1) + AT(x[a], j + 1, k, l) +
                                                      AT(x[a], j + 1, k, l + 1) + AT(x[a], j, k + 1, l) +
                                                      AT(x[a], j + 1, k + 1, l) + AT(x[a], j, k, l + 1));
                                    if (!(c >= io->x1[a] && c <= io->x2[a])) in = 0;
                                }
                                if (!in) continue;
                                if (io->op[0] == '=') AT(f, j, k, l) = val;
                                else if (io->op[0] == '+') AT(f, j, k, l) = AT(f, j, k, l) + val;
                                else AT(f, j, k, l) = AT(f, j, k, l) * val;
                            }
                } else {
                    fprintf(stderr, "oracle: unsupported i/o mode '%s'\n", io->op);
                    exit(3);
                }
                if (io->reg == 'n') scalar_bc(f, bc1, bc2, i1bc, i2bc);
                if (io->reg == 'c') {
                    int t[3] = {i2bc[0] - 1, i2bc[1] - 1, i2bc[2] - 1};
                    scalar_bc(f, bc1, bc2, i1bc, t);
                }
                continue;
            }

            /* spatial indices for disk i/o */
            if (io->op[0] == '.') {
                for (int i = 0; i < 4; i++) m[i] = n[i] = 1, o[i] = 0;
            } else {
                for (int i = 0; i < 3; i++) {
                    i1[i] = io->ii[i][0] - nnoff[i];
                    i2[i] = io->ii[i][1] - nnoff[i];
                    di[i] = io->ii[i][2];
                    i3[i] = i1[i];
                    i4[i] = i2[i];
                    if (i1[i] < i1core[i])
                        i1[i] = i1[i] + (fdiv(i1core[i] - i1[i] - 1, di[i]) + 1) * di[i];
                    if (i2[i] > i2core[i]) i2[i] = i1[i] + fdiv(i2core[i] - i1[i], di[i]) * di[i];
                    m[i] = fdiv(i4[i] - i3[i], di[i]) + 1;
                    n[i] = fdiv(i2[i] - i1[i], di[i]) + 1;
                    o[i] = fdiv(i1[i] - i3[i], di[i]);
                    if (n[i] < 0) n[i] = 0;
                }
                n[3] = 0;
            }
            /* array dimensionality */
            int skip = 0;
            for (int i = 0; i < 3; i++) {
                if (f.n[i] == 1) {
                    if (n[i] < 1) {
                        io->op[0] = '#';
                        skip = 1;
                        break;
                    }
                    i1[i] = i2[i] = 1;
                    m[i] = n[i] = 1;
                    o[i] = 0;
                }
            }
            if (skip) continue;

            /* disk input */
            if (io->op[1] == '<') {
                if (strcmp(io->op, ".<")) {
                    for (int i = 0; i < 3; i++)
                        if (m[i] == 1) {
                            i1[i] = i2[i] = 1;
                            n[i] = 1;
                            o[i] = 0;
                        }
                }
                size_t rec = (size_t)n[0] * n[1] * n[2];
                if (io->ib < 0) {
                    io->buff = ralloc(rec * io->nb);
                    io->ib = io->nb;
                }
                if (io->ib == io->nb) {
                    n[3] = imin(io->nb, fdiv(it2 - it, dit) + 1);
                    m[3] = fdiv(it2 - it1, dit) + 1;
                    o[3] = fdiv(it - it1, dit);
                    if (n[0] >= 1 && n[1] >= 1 && n[2] >= 1 && n[3] >= 1)
                        rio_read(io->fname, io->buff, rec, o[3], n[3]);
                    io->ib = 0;
                    if (n[0] < 1 || n[1] < 1 || n[2] < 1 || n[3] < 1) {
                        free(io->buff);
                        io->op[0] = '#';
                        continue;
                    }
                    for (size_t q2 = 0; q2 < rec * (size_t)n[3]; q2++) {
                        real v = io->buff[q2];
                        if (v != v || v > HUGE_R || v < -HUGE_R) {
                            fprintf(stderr, "NaN/Inf in %s\n", io->fname);
                            exit(1);
                        }
                    }
                }
                io->ib++;
                real *b = io->buff + rec * (size_t)(io->ib - 1);
                if (!strcmp(io->op, ".<")) {
                    /* sub-cell interpolation of the step's value of a point
                       time series (sord/input_output.f90:343-353).  Reference
                       defect: it reads buff(i, ib) with i left over from an
                       earlier loop (:350); the record holds one value, read
                       here as buff(1, ib). */
                    for (int l = i1[2]; l <= i2[2]; l++)
                        for (int k = i1[1]; k <= i2[1]; k++)
                            for (int j = i1[0]; j <= i2[0]; j++) {
                                real r = (R(1.0) - FABS(xi[0] - (real)j)) * (R(1.0) - FABS(xi[1] - (real)k)) *
                                         (R(1.0) - FABS(xi[2] - (real)l));
                                AT(f, j, k, l) = AT(f, j, k, l) + r * b[0];
                            }
                } else {
                size_t c = 0;
                for (int l = i1[2]; l <= i2[2]; l += di[2])
                    for (int k = i1[1]; k <= i2[1]; k += di[1])
                        for (int j = i1[0]; j <= i2[0]; j += di[0]) AT(s1, j, k, l) = b[c++];
                if (di[0] > 1 || di[1] > 1 || di[2] > 1) {
                    int i3[3], i4[3];
                    for (int a = 0; a < 3; a++) i3[a] = io->ii[a][0] - nnoff[a], i4[a] = io->ii[a][1] - nnoff[a];
                    interpolate(s1, i3, i4, di);
                }
                /* half-dimension fill */
                if (m[0] == 1) {
                    i2[0] = f.n[0];
                    for (int l = 1; l <= nm[2]; l++)
                        for (int k = 1; k <= nm[1]; k++)
                            for (int i = 2; i <= i2[0]; i++) AT(s1, i, k, l) = AT(s1, 1, k, l);
                }
                if (m[1] == 1) {
                    i2[1] = f.n[1];
                    for (int l = 1; l <= nm[2]; l++)
                        for (int i = 2; i <= i2[1]; i++)
                            for (int j = 1; j <= nm[0]; j++) AT(s1, j, i, l) = AT(s1, j, 1, l);
                }
                if (m[2] == 1) {
                    i2[2] = f.n[2];
                    for (int i = 2; i <= i2[2]; i++)
                        for (int k = 1; k <= nm[1]; k++)
                            for (int j = 1; j <= nm[0]; j++) AT(s1, j, k, i) = AT(s1, j, k, 1);
                }
                for (int l = i1[2]; l <= i2[2]; l++)
                    for (int k = i1[1]; k <= i2[1]; k++)
                        for (int j = i1[0]; j <= i2[0]; j++) {
                            if (!strcmp(io->op, "=<")) AT(f, j, k, l) = AT(s1, j, k, l);
                            else if (!strcmp(io->op, "+<")) AT(f, j, k, l) = AT(f, j, k, l) + AT(s1, j, k, l);
                            else if (!strcmp(io->op, "*<")) AT(f, j, k, l) = AT(f, j, k, l) * AT(s1, j, k, l);
                            else die(1, "bad i/o op");
                        }
                }
                if (io->reg == 'n') scalar_bc(f, bc1, bc2, i1bc, i2bc);
                if (io->reg == 'c') {
                    int t[3] = {i2bc[0] - 1, i2bc[1] - 1, i2bc[2] - 1};
                    scalar_bc(f, bc1, bc2, i1bc, t);
                }
                if (it == it2) {
                    free(io->buff);
                    io->op[0] = '#';
                }
                continue;
            }

            /* disk output */
            if (io->op[1] == '>') {
                if (!strcmp(io->op, ".>")) {
                    if (i2[0] < i1core[0] || i2[0] > i2core[0] || i2[1] < i1core[1] ||
                        i2[1] > i2core[1] || i2[2] < i1core[2] || i2[2] > i2core[2]) {
                        io->op[0] = '#';
                        continue;
                    }
                }
                size_t rec = (size_t)n[0] * n[1] * n[2];
                if (io->ib < 0) {
                    io->buff = ralloc(rec * io->nb);
                    io->ib = 0;
                    io->fh_open = 0;
                }
                if (fmodulo(it, itstats) != 0) {
                    if (!strcmp(field, "vm2")) vector_norm(f, vv, i1, i2, di);
                    if (!strcmp(field, "um2")) vector_norm(f, uu, i1, i2, di);
                    if (!strcmp(field, "wm2")) tensor_norm(f, w1, w2, i1, i2, di);
                    if (!strcmp(field, "am2")) vector_norm(f, w1, i1, i2, di);
                }
                io->ib++;
                real *b = io->buff + rec * (size_t)(io->ib - 1);
                if (!strcmp(io->op, ".>")) {
                    b[0] = 0;
                    for (int l = i1[2]; l <= i2[2]; l++)
                        for (int k = i1[1]; k <= i2[1]; k++)
                            for (int j = i1[0]; j <= i2[0]; j++) {
                                real r = (R(1.0) - FABS(xi[0] - (real)j)) *
                                         (R(1.0) - FABS(xi[1] - (real)k)) *
                                         (R(1.0) - FABS(xi[2] - (real)l));
                                b[0] = b[0] + AT(f, j, k, l) * r;
                            }
                } else if (!strcmp(io->op, "=>")) {
                    size_t c = 0;
                    for (int l = i1[2]; l <= i2[2]; l += di[2])
                        for (int k = i1[1]; k <= i2[1]; k += di[1])
                            for (int j = i1[0]; j <= i2[0]; j += di[0]) b[c++] = AT(f, j, k, l);
                } else {
                    die(1, "bad i/o op");
                }
                if (io->ib == io->nb || it == it2 || fmodulo(it, itio) == 0) {
                    n[3] = io->ib;
                    m[3] = fdiv(it2 - it1, dit) + 1;
                    o[3] = fdiv(it - it1, dit) + 1 - n[3];
                    if (n[0] >= 1 && n[1] >= 1 && n[2] >= 1 && n[3] >= 1)
                        rio_write(io->fname, io->buff, rec, o[3], n[3]);
                    io->ib = 0;
                    if (it == it2 || n[0] < 1 || n[1] < 1 || n[2] < 1 || n[3] < 1) {
                        free(io->buff);
                        io->op[0] = '#';
                    }
                }
                continue;
            }
            die(1, "bad i/o op");
        }
    }
    timers[1] += clock_ms();
}

/* ------------------------------------------------------------------------ */
/* init_grid: sord/grid_generation.f90:6-200                                */
/* ------------------------------------------------------------------------ */
static real powi_(real x, int n) { /* integer power by repeated squaring */
    real r = 1;
    int neg = n < 0;
    unsigned u = neg ? -n : n;
    while (u) {
        if (u & 1) r *= x;
        x *= x;
        u >>= 1;
    }
    return neg ? R(1.0) / r : r;
}

static void init_grid(void) {
    int i1[3], i2[3], i3[3], i4[3];
    for (size_t q = 0; q < 3 * vol_n; q++) xx[q] = 0;
    for (int a = 0; a < 3; a++) {
        A3 x = comp(xx, a + 1);
        for (int i = i1node[a]; i <= i2node[a]; i++) plane_set(x, a, i, dx[a] * (real)(i + nnoff[a] - 1));
    }
    if (ifn) {
        int a = ifn - 1;
        int lo = imax(i1core[a], irup + 1);
        A3 x = comp(xx, a + 1);
        for (int i = lo; i <= i2node[a]; i++) plane_set(x, a, i, dx[a] * (real)(i + nnoff[a] - 2));
    }
    field_io("<", "x", comp(xx, 1));
    field_io("<", "y", comp(xx, 2));
    field_io("<", "z", comp(xx, 3));
    if (gridnoise > 0) die(3, "oracle: gridnoise (Fortran RNG) is not restated");
    if (rexpand > 1) {
        for (int a = 0; a < 3; a++) {
            i1[a] = n1expand[a] - nnoff[a];
            i2[a] = nn[a] - n2expand[a] + 1 - nnoff[a];
            i3[a] = i1node[a];
            i4[a] = i2node[a];
        }
        for (int a = 0; a < 3; a++) {
            A3 x = comp(xx, a + 1);
            for (int j = i3[a]; j <= imin(i4[a], i1[a]); j++) {
                int i = i1[a] - j;
                real d = dx[a] * ((real)(i + 1) - (powi_(rexpand, i + 1) - R(1.0)) / (rexpand - R(1.0)));
                int lo[3] = {1, 1, 1}, hi[3] = {nm[0], nm[1], nm[2]};
                lo[a] = hi[a] = j;
                for (int l = lo[2]; l <= hi[2]; l++)
                    for (int k = lo[1]; k <= hi[1]; k++)
                        for (int jj = lo[0]; jj <= hi[0]; jj++) AT(x, jj, k, l) = AT(x, jj, k, l) + d;
            }
            for (int j = imax(i3[a], i2[a]); j <= i4[a]; j++) {
                int i = j - i2[a];
                real d = dx[a] * ((real)(i + 1) - (powi_(rexpand, i + 1) - R(1.0)) / (rexpand - R(1.0)));
                int lo[3] = {1, 1, 1}, hi[3] = {nm[0], nm[1], nm[2]};
                lo[a] = hi[a] = j;
                for (int l = lo[2]; l <= hi[2]; l++)
                    for (int k = lo[1]; k <= hi[1]; k++)
                        for (int jj = lo[0]; jj <= hi[0]; jj++) AT(x, jj, k, l) = AT(x, jj, k, l) - d;
            }
        }
    }
    /* affine grid transformation */
    {
        real *m = affine;
        real *X = xx, *Y = xx + vol_n, *Z = xx + 2 * vol_n;
#pragma omp parallel for schedule(static)
        for (size_t q = 0; q < vol_n; q++) {
            real x = m[0] * X[q] + m[1] * Y[q] + m[2] * Z[q];
            real y = m[3] * X[q] + m[4] * Y[q] + m[5] * Z[q];
            real z = m[6] * X[q] + m[7] * Y[q] + m[8] * Z[q];
            X[q] = x;
            Y[q] = y;
            Z[q] = z;
        }
    }
    /* fill halo, bc=4 means copy into halo */
    {
        int bc[3] = {4, 4, 4};
        for (int a = 0; a < 3; a++) {
            i1[a] = i1bc[a] - 1;
            i2[a] = i2bc[a] + 1;
        }
        vector_bc(xx, nm, bc, bc, i1, i2);
    }
    /* cell centers */
    for (int a = 1; a <= 3; a++) {
        average(comp(w1, a), comp(xx, a), i1cell, i2cell, 1);
        set_halo(comp(w1, a), 0, i1cell, i2cell);
    }
    field_io(">", "x", comp(xx, 1));
    field_io(">", "y", comp(xx, 2));
    field_io(">", "z", comp(xx, 3));
    field_io(">", "xc", comp(w1, 1));
    field_io(">", "yc", comp(w1, 2));
    field_io(">", "zc", comp(w1, 3));
    /* cell volume */
    set_halo(vc, 0, i1cell, i2cell);
    int err = 0;
    for (int i = 1; i <= 3; i++) {
        diff_nc(vc, xx, i, i, i1cell, i2cell, xx);
        if (ifn) plane_set(vc, ifn - 1, irup, 0);
        real mn = vc.p[0];
        for (size_t q = 0; q < vol_n; q++)
            if (vc.p[q] < mn) mn = vc.p[q];
        err = mn < 0;
    }
    field_io(">", "vc", vc);
    if (err) die(1, "Negative cell volume. Wrong sign in dx or problem with mesh.");
}

/* ------------------------------------------------------------------------ */
/* init_material: sord/material_model.f90:6-171                             */
/* ------------------------------------------------------------------------ */
static void init_material(void) {
    real maxl[14], vmin_, vmax_, cfl1, cfl2, r, s;
    for (size_t q = 0; q < vol_n; q++) mr.p[q] = lam.p[q] = mu.p[q] = gam.p[q] = 0;
    field_io("<", "rho", mr);
    field_io("<", "vp", lam);
    field_io("<", "vs", mu);
    field_io("<", "gam", gam);
    memcpy(s1.p, lam.p, vol_n * sizeof(real));
    memcpy(s2.p, mu.p, vol_n * sizeof(real));
    rlimits(mr.p, rho_min, rho_max, vol_n);
    rlimits(s1.p, vp_min, vp_max, vol_n);
    rlimits(s2.p, vs_min, vs_max, vol_n);
    if (vdamp > 0)
        for (size_t q = 0; q < vol_n; q++) gam.p[q] = s2.p[q] == 0 ? 0 : vdamp / s2.p[q];
    rlimits(gam.p, gam_min, gam_max, vol_n);
    set_halo(mr, 0, i1cell, i2cell);
    set_halo(s1, 0, i1cell, i2cell);
    set_halo(s2, 0, i1cell, i2cell);
    set_halo(gam, 0, i1cell, i2cell);
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < vol_n; q++) {
        mu.p[q] = mr.p[q] * s2.p[q] * s2.p[q];
        lam.p[q] = mr.p[q] * s1.p[q] * s1.p[q] - R(2.0) * mu.p[q];
        real rr = lam.p[q] + mu.p[q];
        yy.p[q] = rr == 0 ? 0 : R(0.5) * lam.p[q] / rr;
    }
    int i1[3], i2[3];
    for (int a = 0; a < 3; a++) {
        i1[a] = imax(i1core[a], i1bc[a]);
        i2[a] = imin(i2core[a], i2bc[a] - 1);
    }
    A3 arrs[7] = {mr, s1, s2, gam, lam, mu, yy};
    for (int q = 0; q < 7; q++) {
        real mn = HUGE_R, mx = -HUGE_R;
        for (int l = i1[2]; l <= i2[2]; l++)
            for (int k = i1[1]; k <= i2[1]; k++)
                for (int j = i1[0]; j <= i2[0]; j++) {
                    real v = AT(arrs[q], j, k, l);
                    if (v < mn) mn = v;
                    if (v > mx) mx = v;
                }
        maxl[q] = -mn;
        maxl[q + 7] = mx;
    }
    field_io(">", "rho", mr);
    field_io(">", "vp", s1);
    field_io(">", "vs", s2);
    field_io(">", "gam", gam);
    field_io(">", "lam", lam);
    field_io(">", "mu", mu);
    field_io(">", "nu", yy);
    s = SQRT((dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]) / R(3.0)) / R(12.0);
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < vol_n; q++) {
        real rr = lam.p[q] + R(2.0) * mu.p[q];
        yy.p[q] = rr == 0 ? 0 : mu.p[q] * (lam.p[q] + mu.p[q]) * s / rr;
    }
    real *maxg = maxl; /* rreduce1 allmax: identity in the serial module */
    if (vpml <= 0) {
        vmin_ = -maxg[2];
        vmax_ = maxg[8];
        vpml = R(2.0) * vmin_ * vmax_ / (vmin_ + vmax_);
    }
    vmax_ = maxg[8];
    r = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
    cfl1 = dt * vmax_ * SQRT(R(3.0) / r);
    cfl2 = dt * vmax_ * R(3.0) / SQRT(r);
    FILE *fp = fopen("material.json", "w");
    if (fp) {
        fprintf(fp, "{\n    \"cfl\": [%.9g, %.9g],\n", (double)cfl1, (double)cfl2);
        const char *nm_[7] = {"rho", "vp", "vs", "gam", "lam", "mu", "nu"};
        for (int q = 0; q < 7; q++)
            fprintf(fp, "    \"%s\": [%.9g, %.9g]%s\n", nm_[q], (double)-maxg[q], (double)maxg[q + 7],
                    q < 6 ? "," : "");
        fprintf(fp, "}\n");
        fclose(fp);
    }
    for (int q = 0; q < 7; q++)
        if (maxg[q] > 0) die(1, "negative moduli");
    if (cfl2 > R(1.0)) die(1, "Courant condition not satisfied!");
}

/* ------------------------------------------------------------------------ */
/* init_pml: sord/boundary_pml.f90:7-82                                     */
/* ------------------------------------------------------------------------ */
static void init_pml(void) {
    if (npml < 1) return;
    real c1 = R(8.0) / R(15.0), c2 = R(-3.0) / R(100.0), c3 = R(1.0) / R(1500.0), tune = R(3.5);
    real r = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
    int np = 1;
    for (int q = 0; q < ppml; q++) np *= npml;
    real damp = tune * vpml / SQRT(r / R(3.0)) * (c1 + (c2 + c3 * (real)npml) * (real)npml) / (real)np;
    dn1 = ralloc(npml);
    dn2 = ralloc(npml);
    dc1 = ralloc(npml);
    dc2 = ralloc(npml);
    for (int i = 1; i <= npml; i++) {
        int ip = 1, im = 1;
        for (int q = 0; q < ppml; q++) ip *= i, im *= (i - 1);
        real dampn = damp * (real)ip;
        real dampc = damp * (real)(ip + im) * R(0.5);
        dn1[npml - i] = -(R(2.0) * dampn) / (R(2.0) + dt * dampn);
        dc1[npml - i] = (R(2.0) - dt * dampc) / (R(2.0) + dt * dampc);
        dn2[npml - i] = R(2.0) / (R(2.0) + dt * dampn);
        dc2[npml - i] = R(2.0) * dt / (R(2.0) + dt * dampc);
    }
    int i1[3], i2[3];
    for (int a = 0; a < 3; a++) {
        i1[a] = imax(imin(i2node[a], i1pml[a]) - i1node[a] + 1, 0);
        i2[a] = imax(i2node[a] - imax(i1node[a], i2pml[a]) + 1, 0);
    }
    for (int a = 0; a < 3; a++) {
        for (int q = 0; q < 3; q++) pn[a][q] = pn[a + 3][q] = nm[q];
        pn[a][a] = i1[a];
        pn[a + 3][a] = i2[a];
    }
    for (int a = 0; a < 3; a++) {
        i1[a] = imax(imin(i2cell[a], i1pml[a]) - i1cell[a] + 1, 0);
        i2[a] = imax(i2cell[a] - imax(i1cell[a], i2pml[a] - 1) + 1, 0);
    }
    for (int a = 0; a < 3; a++) {
        for (int q = 0; q < 3; q++) gn[a][q] = gn[a + 3][q] = nm[q];
        gn[a][a] = i1[a];
        gn[a + 3][a] = i2[a];
    }
    real **pp[6] = {&p1, &p2, &p3, &p4, &p5, &p6}, **gg[6] = {&g1, &g2, &g3, &g4, &g5, &g6};
    for (int s = 0; s < 6; s++) {
        *pp[s] = ralloc((size_t)pn[s][0] * pn[s][1] * pn[s][2] * 3);
        *gg[s] = ralloc((size_t)gn[s][0] * gn[s][1] * gn[s][2] * 3);
    }
}
/* view of component ic of PML slab s */
static A3 slab(real *p, int s, int ic, int isg) {
    int *n = isg ? gn[s] : pn[s];
    A3 a = {p + (size_t)n[0] * n[1] * n[2] * (size_t)(ic - 1), {n[0], n[1], n[2]}};
    return a;
}

/* ------------------------------------------------------------------------ */
/* init_rupture: sord/dynamic_rupture.f90:7-238                             */
/* ------------------------------------------------------------------------ */
static void init_rupture(void) {
    if (!strcmp(faultnormal, "no")) return;
    if (ifn == 0) return;
    for (int a = 0; a < 3; a++) fm[a] = nm[a];
    fm[ifn - 1] = 1;
    size_t nf = (size_t)fm[0] * fm[1] * fm[2];
    A3 *sc[16] = {&mus, &mud, &dc, &co, &area, &rhypo, &lamf, &muf, &sl, &psv, &trup, &tarr, &tn, &ts, &f1, &f2};
    for (int q = 0; q < 16; q++) {
        sc[q]->p = ralloc(nf);
        sc[q]->n[0] = fm[0];
        sc[q]->n[1] = fm[1];
        sc[q]->n[2] = fm[2];
    }
    nhat = ralloc(3 * nf);
    t0 = ralloc(3 * nf);
    t1 = ralloc(3 * nf);
    t2 = ralloc(3 * nf);
    t3 = ralloc(3 * nf);
    field_io("<>", "mus", mus);
    field_io("<>", "mud", mud);
    field_io("<>", "dc", dc);
    field_io("<>", "co", co);
    field_io("<>", "sxx", fcomp(t1, 1));
    field_io("<>", "syy", fcomp(t1, 2));
    field_io("<>", "szz", fcomp(t1, 3));
    field_io("<>", "syz", fcomp(t2, 1));
    field_io("<>", "szx", fcomp(t2, 2));
    field_io("<>", "sxy", fcomp(t2, 3));
    field_io("<>", "ts", fcomp(t3, 1));
    field_io("<>", "td", fcomp(t3, 2));
    field_io("<>", "tn", fcomp(t3, 3));
    {
        real mx = -HUGE_R;
        for (size_t q = 0; q < nf; q++)
            if (t3[2 * nf + q] > mx) mx = t3[2 * nf + q];
        if (mx > 0) fprintf(stderr, "Warning: positive normal traction: %g\n", (double)mx);
    }
    int i1[3], i2[3];
    for (int a = 0; a < 3; a++) {
        i1[a] = i1pml[a] + 1;
        i2[a] = i2pml[a] - 1;
    }
    set_halo(co, R(1e20), i1, i2);
    for (int a = 0; a < 3; a++) {
        i1[a] = i1core[a];
        i2[a] = i2core[a];
    }
    i1[ifn - 1] = i2[ifn - 1] = irup;
    node_normals(nhat, w1, i1, i2, ifn);
    for (size_t q = 0; q < nf; q++) {
        real s = SQRT(nhat[q] * nhat[q] + nhat[nf + q] * nhat[nf + q] + nhat[2 * nf + q] * nhat[2 * nf + q]);
        area.p[q] = faultnormal[0] == '+' ? s : -s;
        f1.p[q] = area.p[q];
    }
    rinvert(f1.p, nf);
    for (int i = 0; i < 3; i++)
        for (size_t q = 0; q < nf; q++) nhat[i * nf + q] = nhat[i * nf + q] * f1.p[q];
    field_io(">", "nsx", fcomp(nhat, 1));
    field_io(">", "nsy", fcomp(nhat, 2));
    field_io(">", "nsz", fcomp(nhat, 3));
    /* resolve prestress onto fault */
    for (int i = 1; i <= 3; i++) {
        int j = i % 3 + 1, k = (i + 1) % 3 + 1;
        for (size_t q = 0; q < nf; q++)
            t0[(i - 1) * nf + q] = t1[(i - 1) * nf + q] * nhat[(i - 1) * nf + q] +
                                   t2[(j - 1) * nf + q] * nhat[(k - 1) * nf + q] +
                                   t2[(k - 1) * nf + q] * nhat[(j - 1) * nf + q];
    }
    /* Ts2 vector */
    for (size_t q = 0; q < nf; q++) {
        real n1 = nhat[q], n2 = nhat[nf + q], n3 = nhat[2 * nf + q];
        t2[q] = n2 * slipvector[2] - n3 * slipvector[1];
        t2[nf + q] = n3 * slipvector[0] - n1 * slipvector[2];
        t2[2 * nf + q] = n1 * slipvector[1] - n2 * slipvector[0];
        f1.p[q] = SQRT(t2[q] * t2[q] + t2[nf + q] * t2[nf + q] + t2[2 * nf + q] * t2[2 * nf + q]);
    }
    rinvert(f1.p, nf);
    for (int i = 0; i < 3; i++)
        for (size_t q = 0; q < nf; q++) t2[i * nf + q] = t2[i * nf + q] * f1.p[q];
    /* Ts1 vector */
    for (size_t q = 0; q < nf; q++) {
        real n1 = nhat[q], n2 = nhat[nf + q], n3 = nhat[2 * nf + q];
        t1[q] = t2[nf + q] * n3 - t2[2 * nf + q] * n2;
        t1[nf + q] = t2[2 * nf + q] * n1 - t2[q] * n3;
        t1[2 * nf + q] = t2[q] * n2 - t2[nf + q] * n1;
        f1.p[q] = SQRT(t1[q] * t1[q] + t1[nf + q] * t1[nf + q] + t1[2 * nf + q] * t1[2 * nf + q]);
    }
    rinvert(f1.p, nf);
    for (int i = 0; i < 3; i++)
        for (size_t q = 0; q < nf; q++) t1[i * nf + q] = t1[i * nf + q] * f1.p[q];
    /* total pretraction */
    for (int i = 0; i < 3; i++)
        for (size_t q = 0; q < nf; q++)
            t0[i * nf + q] = t0[i * nf + q] + t3[q] * t1[i * nf + q] + t3[nf + q] * t2[i * nf + q] +
                             t3[2 * nf + q] * nhat[i * nf + q];
    /* hypocentral radius */
    if (rcrit > 0 && vrup > 0) {
        real xhypo[3] = {0, 0, 0}, xi[3];
        int ii[3], inside = 1;
        for (int a = 0; a < 3; a++) {
            xi[a] = hypocenter[a] - (real)nnoff[a] + R(1.0);
            ii[a] = (int)floor((double)xi[a]);
            if (!(ii[a] >= 1 && ii[a] < nm[a])) inside = 0;
        }
        if (inside) {
            for (int l = ii[2]; l <= ii[2] + 1; l++)
                for (int k = ii[1]; k <= ii[1] + 1; k++)
                    for (int j = ii[0]; j <= ii[0] + 1; j++) {
                        real w = (R(1.0) - FABS(xi[0] - (real)j)) * (R(1.0) - FABS(xi[1] - (real)k)) *
                                 (R(1.0) - FABS(xi[2] - (real)l));
                        for (int i = 1; i <= 3; i++) xhypo[i - 1] = xhypo[i - 1] + w * AT(comp(w1, i), j, k, l);
                    }
        }
        for (int i = 1; i <= 3; i++) {
            A3 t = fcomp(t2, i), w = comp(w1, i);
            for (int l = 1; l <= fm[2]; l++)
                for (int k = 1; k <= fm[1]; k++)
                    for (int j = 1; j <= fm[0]; j++) {
                        int jj = j, kk = k, ll = l;
                        if (ifn == 1) jj = irup;
                        if (ifn == 2) kk = irup;
                        if (ifn == 3) ll = irup;
                        AT(t, j, k, l) = AT(w, jj, kk, ll) - xhypo[i - 1];
                    }
        }
        for (size_t q = 0; q < nf; q++)
            rhypo.p[q] = SQRT(t2[q] * t2[q] + t2[nf + q] * t2[nf + q] + t2[2 * nf + q] * t2[2 * nf + q]);
    }
    /* moduli on the fault plane */
    for (int l = 1; l <= fm[2]; l++)
        for (int k = 1; k <= fm[1]; k++)
            for (int j = 1; j <= fm[0]; j++) {
                int jj = j, kk = k, ll = l;
                if (ifn == 1) jj = irup;
                if (ifn == 2) kk = irup;
                if (ifn == 3) ll = irup;
                AT(lamf, j, k, l) = AT(lam, jj, kk, ll);
                AT(muf, j, k, l) = AT(mu, jj, kk, ll);
            }
    rinvert(lamf.p, nf);
    rinvert(muf.p, nf);
    A3 fa[2] = {lamf, muf};
    for (int q = 0; q < 2; q++) {
        A3 f = fa[q];
        real *tmp = ralloc(nf);
        for (int a = 0; a < 3; a++) {
            if (ifn == a + 1) continue;
            memcpy(tmp, f.p, nf * sizeof(real));
            A3 t = {tmp, {fm[0], fm[1], fm[2]}};
            int hi = nm[a] - 1;
            for (int l = 1; l <= fm[2]; l++)
                for (int k = 1; k <= fm[1]; k++)
                    for (int j = 1; j <= fm[0]; j++) {
                        int c = a == 0 ? j : a == 1 ? k : l;
                        if (c < 2 || c > hi) continue;
                        int jm = j - (a == 0), km = k - (a == 1), lm = l - (a == 2);
                        AT(f, j, k, l) = R(0.5) * (AT(t, j, k, l) + AT(t, jm, km, lm));
                    }
        }
        free(tmp);
    }
    rinvert(muf.p, nf);
    for (size_t q = 0; q < nf; q++) {
        psv.p[q] = 0;
        trup.p[q] = R(1e9);
        tarr.p[q] = 0;
    }
    efric = 0;
}

/* ------------------------------------------------------------------------ */
/* resample_material: sord/material_resample.f90:6-55                       */
/* ------------------------------------------------------------------------ */
static void resample_material(void) {
    int i1[3], i2[3], bc[3] = {4, 4, 4};
    for (size_t q = 0; q < vol_n; q++) s2.p[q] = mr.p[q] * vc.p[q];
    average(mr, s2, i1node, i2node, -1);
    rinvert(mr.p, vol_n);
    scalar_bc(mr, bc1, bc2, i1bc, i2bc);
    rinvert(vc.p, vol_n);
    for (int a = 0; a < 3; a++) {
        i1[a] = i1bc[a] - 1;
        i2[a] = i2bc[a];
    }
    scalar_bc(gam, bc, bc, i1, i2);
    for (size_t q = 0; q < vol_n; q++) s2.p[q] = gam.p[q] * dt;
    average(gam, s2, i1node, i2node, -1);
    set_halo(gam, 0, i1bc, i2bc);
    scalar_bc(gam, bc1, bc2, i1bc, i2bc);
    for (int a = 0; a < 3; a++) {
        i1[a] = i1bc[a];
        i2[a] = i2bc[a] - 1;
    }
    set_halo(yy, 0, i1, i2);
    if (ifn) plane_set(yy, ifn - 1, irup, 0);
    memset(vv, 0, 3 * vol_n * sizeof(real));
    memset(uu, 0, 3 * vol_n * sizeof(real));
    memset(w1, 0, 3 * vol_n * sizeof(real));
    memset(w2, 0, 3 * vol_n * sizeof(real));
    memset(s1.p, 0, vol_n * sizeof(real));
    memset(s2.p, 0, vol_n * sizeof(real));
}

/* maxloc over a whole array (first occurrence, Fortran order) */
static real maxloc_(A3 f, int loc[3]) {
    size_t best = 0;
    size_t n = (size_t)f.n[0] * f.n[1] * f.n[2];
    for (size_t q = 1; q < n; q++)
        if (f.p[q] > f.p[best]) best = q;
    loc[0] = (int)(best % f.n[0]) + 1;
    loc[1] = (int)((best / f.n[0]) % f.n[1]) + 1;
    loc[2] = (int)(best / ((size_t)f.n[0] * f.n[1])) + 1;
    return f.p[best];
}
static real maxval_(A3 f) {
    int loc[3];
    return maxloc_(f, loc);
}
static real minval_(A3 f) {
    size_t n = (size_t)f.n[0] * f.n[1] * f.n[2];
    real m = f.p[0];
    for (size_t q = 1; q < n; q++)
        if (f.p[q] < m) m = f.p[q];
    return m;
}
static real sum_(A3 f) {
    size_t n = (size_t)f.n[0] * f.n[1] * f.n[2];
    real s = 0;
    for (size_t q = 0; q < n; q++) s = s + f.p[q];
    return s;
}
/* fault-plane difference of a volume vector: t = v(irup+1) - v(irup) */
static void plane_diff(real *t, real *v) {
    for (int i = 1; i <= 3; i++) {
        A3 a = fcomp(t, i), b = comp(v, i);
        for (int l = 1; l <= fm[2]; l++)
            for (int k = 1; k <= fm[1]; k++)
                for (int j = 1; j <= fm[0]; j++) {
                    int jj = j, kk = k, ll = l, jp = j, kp = k, lp = l;
                    if (ifn == 1) jj = irup, jp = irup + 1;
                    if (ifn == 2) kk = irup, kp = irup + 1;
                    if (ifn == 3) ll = irup, lp = irup + 1;
                    AT(a, j, k, l) = AT(b, jp, kp, lp) - AT(b, jj, kk, ll);
                }
    }
}
static void vnorm3(A3 f, real *t) { /* f = sqrt(sum(t*t, 4)) on the fault plane */
    size_t nf = (size_t)fm[0] * fm[1] * fm[2];
    for (size_t q = 0; q < nf; q++) f.p[q] = SQRT(t[q] * t[q] + t[nf + q] * t[nf + q] + t[2 * nf + q] * t[2 * nf + q]);
}

/* ------------------------------------------------------------------------ */
/* step_time: sord/time_integration.f90:6-115                               */
/* ------------------------------------------------------------------------ */
/* ---- finite kinematic source: sord/kinematic_source.f90 ---- */
static int src_n = 0, *src_nt;
static real *src_xi, *src_dt, *src_t0, *src_w1, *src_w2, *src_history; /* [isrc + n * comp] */

static void iio_read(const char *fname, int *buf, int n) { /* sord/fortran_io.f90 iio1 (int32) */
    FILE *fp = fopen(fname, "rb");
    if (!fp) {
        fprintf(stderr, "cannot open input file %s\n", fname);
        exit(2);
    }
    if (fread(buf, sizeof(int), (size_t)n, fp) != (size_t)n) die(2, "short read");
    fclose(fp);
}

/* sord/kinematic_source.f90:10-64 */
static void init_finite_source(void) {
    if (nsource == 0) return;
    int n = abs(nsource);
    src_xi = ralloc((size_t)n * 3);
    rio_read("source-xi1.bin", src_xi, (size_t)n, 0, 1);
    rio_read("source-xi2.bin", src_xi + n, (size_t)n, 0, 1);
    rio_read("source-xi3.bin", src_xi + 2 * n, (size_t)n, 0, 1);
    const int force = !strcmp(source, "force");
    for (int i = 0; i < 3; i++) {
        int below = 1, above = 1;
        for (int q = 0; q < n; q++) {
            real x = src_xi[q + n * i];
            x = force ? x - (real)nnoff[i] : x - R(0.5) - (real)nnoff[i];
            src_xi[q + n * i] = x;
            if (!(x < R(-1.0) + (real)(force ? i1node[i] : i1cell[i]))) below = 0;
            if (!(x > R(1.0) + (real)(force ? i2node[i] : i2cell[i]))) above = 0;
        }
        if (below || above) {
            nsource = 0;
            return;
        }
    }
    src_w1 = ralloc((size_t)n * 3);
    static const char *f1[3] = {"source-w11.bin", "source-w12.bin", "source-w13.bin"};
    static const char *t1[3] = {"source-w11.bin", "source-w22.bin", "source-w33.bin"};
    static const char *t2[3] = {"source-w23.bin", "source-w31.bin", "source-w12.bin"};
    for (int i = 0; i < 3; i++) rio_read(force ? f1[i] : t1[i], src_w1 + n * i, (size_t)n, 0, 1);
    if (!force) {
        src_w2 = ralloc((size_t)n * 3);
        for (int i = 0; i < 3; i++) rio_read(t2[i], src_w2 + n * i, (size_t)n, 0, 1);
    }
    src_t0 = ralloc((size_t)n);
    src_dt = ralloc((size_t)n);
    src_nt = (int *)malloc(sizeof(int) * (size_t)n);
    rio_read("source-t0.bin", src_t0, (size_t)n, 0, 1);
    rio_read("source-dt.bin", src_dt, (size_t)n, 0, 1);
    iio_read("source-nt.bin", src_nt, n);
    size_t nh = 0;
    for (int q = 0; q < n; q++) nh += (size_t)src_nt[q];
    src_history = ralloc(nh);
    rio_read("source-history.bin", src_history, nh, 0, 1);
    src_n = n;
}

/* sord/kinematic_source.f90:67-118: add the source to the force (w1) or to
 * the potency / moment tensor (w1, w2) */
static void finite_source(void) {
    if (nsource == 0) return;
    const int n = src_n, force = !strcmp(source, "force");
    size_t itoff = 0;
    for (int isrc = 0; isrc < n; isrc++) {
        int i = (int)floor((double)((tm - src_t0[isrc]) / src_dt[isrc])) + 1;
        real xi[3];
        int i1[3], i2[3], ok = 1;
        for (int a = 0; a < 3; a++) {
            xi[a] = src_xi[isrc + n * a];
            int fl = (int)floor((double)xi[a]);
            i1[a] = imax(force ? i1node[a] : i1cell[a], fl);
            i2[a] = imin(force ? i2node[a] : i2cell[a], fl + 1);
            if (i2[a] < i1[a]) ok = 0;
        }
        if (i >= 0 && ok) {
            i = imin(i, src_nt[isrc] - 1);
            real t = src_t0[isrc] + src_dt[isrc] * (real)(i - 1);
            real h = fminr(R(1.0), (tm - t) / src_dt[isrc]);
            if (i == 0) h = h * src_history[itoff];
            else h = (R(1.0) - h) * src_history[itoff + (size_t)i - 1] + h * src_history[itoff + (size_t)i];
            for (int l = i1[2]; l <= i2[2]; l++)
                for (int k = i1[1]; k <= i2[1]; k++)
                    for (int j = i1[0]; j <= i2[0]; j++) {
                        real w = h * ((R(1.0) - FABS(xi[0] - (real)j)) * (R(1.0) - FABS(xi[1] - (real)k)) *
                                      (R(1.0) - FABS(xi[2] - (real)l)));
                        for (int c = 1; c <= 3; c++) {
                            A3 a = comp(w1, c);
                            if (force) {
                                AT(a, j, k, l) = AT(a, j, k, l) + w * src_w1[isrc + n * (c - 1)];
                            } else {
                                A3 b = comp(w2, c);
                                AT(a, j, k, l) = AT(a, j, k, l) - w * src_w1[isrc + n * (c - 1)];
                                AT(b, j, k, l) = AT(b, j, k, l) - w * src_w2[isrc + n * (c - 1)];
                            }
                        }
                    }
        }
        itoff += (size_t)src_nt[isrc];
    }
}

static void step_time(void) {
    size_t nf = (size_t)fm[0] * fm[1] * fm[2];
    int one[3] = {1, 1, 1};
    if (ifn) {
        plane_diff(t2, vv);
        vnorm3(f2, t2);
    }
    tm = tm0 + dt * (real)(it - 1);
    for (int i = 0; i < 3; i++) {
        real *V = vv + i * vol_n, *U = uu + i * vol_n, *W = w1 + i * vol_n;
#pragma omp parallel for schedule(static)
        for (size_t q = 0; q < vol_n; q++) {
            real v = V[q] + dt * W[q];
            real u = U[q] + dt * v;
            V[q] = v;
            U[q] = u;
            W[q] = u + gam.p[q] * v;
        }
    }
    field_io("<>", "vx", comp(vv, 1));
    field_io("<>", "vy", comp(vv, 2));
    field_io("<>", "vz", comp(vv, 3));
    if (fmodulo(it, itstats) == 0) {
        vector_norm(s1, vv, i1core, i2core, one);
        set_halo(s1, -1, i1core, i2core);
        vmax = maxloc_(s1, vmaxloc);
    }
    field_io(">", "vm2", s1);
    field_io("<>", "ux", comp(uu, 1));
    field_io("<>", "uy", comp(uu, 2));
    field_io("<>", "uz", comp(uu, 3));
    if (fmodulo(it, itstats) == 0) {
        vector_norm(s1, uu, i1core, i2core, one);
        set_halo(s1, -1, i1core, i2core);
        umax = maxloc_(s1, umaxloc);
    }
    field_io(">", "um2", s1);
    if (ifn) {
        plane_diff(t1, vv);
        vnorm3(f1, t1);
        for (size_t q = 0; q < nf; q++) {
            sl.p[q] = sl.p[q] + dt * f1.p[q];
            psv.p[q] = fmaxr(psv.p[q], f1.p[q]);
        }
        if (svtol > 0) {
            for (size_t q = 0; q < nf; q++)
                if (f1.p[q] >= svtol && trup.p[q] > R(1e8))
                    trup.p[q] = tm - dt * (R(0.5) + (svtol - f1.p[q]) / (f2.p[q] - f1.p[q]));
            for (size_t q = 0; q < nf; q++)
                if (f1.p[q] >= svtol) tarr.p[q] = R(1e9);
            for (size_t q = 0; q < nf; q++)
                if (f1.p[q] < svtol && f2.p[q] >= svtol)
                    tarr.p[q] = tm - dt * (R(0.5) + (svtol - f1.p[q]) / (f2.p[q] - f1.p[q]));
        }
        plane_diff(t2, uu);
        vnorm3(f2, t2);
        field_io(">", "svx", fcomp(t1, 1));
        field_io(">", "svy", fcomp(t1, 2));
        field_io(">", "svz", fcomp(t1, 3));
        field_io(">", "svm", f1);
        field_io(">", "psv", psv);
        field_io(">", "sux", fcomp(t2, 1));
        field_io(">", "suy", fcomp(t2, 2));
        field_io(">", "suz", fcomp(t2, 3));
        field_io(">", "sum", f2);
        field_io(">", "sl", sl);
        field_io(">", "trup", trup);
        field_io(">", "tarr", tarr);
        set_halo(f1, -1, i1core, i2core);
        set_halo(f2, -1, i1core, i2core);
        set_halo(tarr, -1, i1core, i2core);
        svmax = maxval_(f1);
        sumax = maxval_(f2);
        slmax = maxval_(sl);
        tarrmax = maxval_(tarr);
    }
}

/* ------------------------------------------------------------------------ */
/* step_stress: sord/stress.f90:6-255                                       */
/* ------------------------------------------------------------------------ */
static void step_stress(void) {
    int i1[3], i2[3], one[3] = {1, 1, 1};
    set_halo(s1, 0, i1cell, i2cell);
    for (int ic = 1; ic <= 3; ic++)
        for (int iid = 1; iid <= 3; iid++) {
            int id = fmodulo(ic + iid - 1, 3) + 1;
            /* elastic region: g_ij = (u_i + gamma*v_i),j */
            for (int a = 0; a < 3; a++) {
                i1[a] = imax(i1pml[a] + 1, i1cell[a]);
                i2[a] = imin(i2pml[a] - 2, i2cell[a]);
            }
            diff_nc(s1, w1, ic, id, i1, i2, xx);
            /* pml region, non-damped directions: g_ij = u_i,j */
            for (int i = 1; i <= 3; i++) {
                if (id == i) continue;
                for (int a = 0; a < 3; a++) i1[a] = i1cell[a], i2[a] = i2cell[a];
                i2[i - 1] = imin(i2[i - 1], i1pml[i - 1]);
                diff_nc(s1, uu, ic, id, i1, i2, xx);
                for (int a = 0; a < 3; a++) i1[a] = i1cell[a], i2[a] = i2cell[a];
                i1[i - 1] = imax(i1[i - 1], i2pml[i - 1] - 1);
                diff_nc(s1, uu, ic, id, i1, i2, xx);
            }
            /* pml region, damped direction: g'_ij = d_j*g_ij = v_i,j */
            {
                int a = id - 1;
                for (int b = 0; b < 3; b++) i1[b] = i1cell[b], i2[b] = i2cell[b];
                i2[a] = imin(i2[a], i1pml[a]);
                diff_nc(s1, vv, ic, id, i1, i2, xx);
                A3 g = slab(a == 0 ? g1 : a == 1 ? g2 : g3, a, ic, 1);
                for (int l = i1[2]; l <= i2[2]; l++)
                    for (int k = i1[1]; k <= i2[1]; k++)
                        for (int j = i1[0]; j <= i2[0]; j++) {
                            int c = a == 0 ? j : a == 1 ? k : l;
                            int i = c - i1[a] + 1, p = c + nnoff[a];
                            int gj = a == 0 ? i : j, gk = a == 1 ? i : k, gl = a == 2 ? i : l;
                            AT(s1, j, k, l) = dc2[p - 1] * AT(s1, j, k, l) + dc1[p - 1] * AT(g, gj, gk, gl);
                            AT(g, gj, gk, gl) = AT(s1, j, k, l);
                        }
                for (int b = 0; b < 3; b++) i1[b] = i1cell[b], i2[b] = i2cell[b];
                i1[a] = imax(i1[a], i2pml[a] - 1);
                diff_nc(s1, vv, ic, id, i1, i2, xx);
                g = slab(a == 0 ? g4 : a == 1 ? g5 : g6, a + 3, ic, 1);
                for (int l = i1[2]; l <= i2[2]; l++)
                    for (int k = i1[1]; k <= i2[1]; k++)
                        for (int j = i1[0]; j <= i2[0]; j++) {
                            int c = a == 0 ? j : a == 1 ? k : l;
                            int i = i2[a] - c + 1, p = nn[a] - c - nnoff[a];
                            int gj = a == 0 ? i : j, gk = a == 1 ? i : k, gl = a == 2 ? i : l;
                            AT(s1, j, k, l) = dc2[p - 1] * AT(s1, j, k, l) + dc1[p - 1] * AT(g, gj, gk, gl);
                            AT(g, gj, gk, gl) = AT(s1, j, k, l);
                        }
            }
            /* add contribution to strain */
            int i = 6 - ic - id;
            if (ic < id) {
                real *W = w2 + (i - 1) * vol_n;
#pragma omp parallel for schedule(static)
                for (size_t q = 0; q < vol_n; q++) W[q] = R(0.5) * s1.p[q];
            } else if (ic > id) {
                real *W = w2 + (i - 1) * vol_n;
#pragma omp parallel for schedule(static)
                for (size_t q = 0; q < vol_n; q++) W[q] = W[q] + R(0.5) * s1.p[q];
            } else {
                memcpy(w1 + (ic - 1) * vol_n, s1.p, vol_n * sizeof(real));
            }
        }
    if (!strcmp(source, "potency")) finite_source();
    field_io("<>", "pxx", comp(w1, 1));
    field_io("<>", "pyy", comp(w1, 2));
    field_io("<>", "pzz", comp(w1, 3));
    field_io("<>", "pyz", comp(w2, 1));
    field_io("<>", "pzx", comp(w2, 2));
    field_io("<>", "pxy", comp(w2, 3));
    /* Hooke's law */
    {
        real *a = w1, *b = w1 + vol_n, *c = w1 + 2 * vol_n;
#pragma omp parallel for schedule(static)
        for (size_t q = 0; q < vol_n; q++) s1.p[q] = lam.p[q] * (a[q] + b[q] + c[q]);
    }
    for (int i = 0; i < 3; i++) {
        real *a = w1 + i * vol_n, *b = w2 + i * vol_n;
#pragma omp parallel for schedule(static)
        for (size_t q = 0; q < vol_n; q++) {
            a[q] = R(2.0) * mu.p[q] * a[q] + s1.p[q];
            b[q] = R(2.0) * mu.p[q] * b[q];
        }
    }
    if (!strcmp(source, "moment")) finite_source();
    field_io("<>", "mxx", comp(w1, 1));
    field_io("<>", "myy", comp(w1, 2));
    field_io("<>", "mzz", comp(w1, 3));
    field_io("<>", "myz", comp(w2, 1));
    field_io("<>", "mzx", comp(w2, 2));
    field_io("<>", "mxy", comp(w2, 3));
    for (int i = 0; i < 3; i++) {
        real *a = w1 + i * vol_n, *b = w2 + i * vol_n;
#pragma omp parallel for schedule(static)
        for (size_t q = 0; q < vol_n; q++) {
            a[q] = a[q] * vc.p[q];
            b[q] = b[q] * vc.p[q];
        }
    }
    field_io("<>", "wxx", comp(w1, 1));
    field_io("<>", "wyy", comp(w1, 2));
    field_io("<>", "wzz", comp(w1, 3));
    field_io("<>", "wyz", comp(w2, 1));
    field_io("<>", "wzx", comp(w2, 2));
    field_io("<>", "wxy", comp(w2, 3));
    if (fmodulo(it, itstats) == 0) {
        tensor_norm(s1, w1, w2, i1core, i2core, one);
        set_halo(s1, -1, i1core, i2core);
        wmax = maxloc_(s1, wmaxloc);
    }
    field_io(">", "wm2", s1);
}

/* ------------------------------------------------------------------------ */
/* step_rupture: sord/dynamic_rupture.f90:243-384                           */
/* ------------------------------------------------------------------------ */
static void step_rupture(void) {
    if (ifn == 0) return;
    size_t nf = (size_t)fm[0] * fm[1] * fm[2];
    int a = ifn - 1;
    /* gather the two sides: - at irup, + at irup+1 */
#define SIDE(arr4, i, q, plus) AT(comp(arr4, i), qj[q] + ((plus) && a == 0), qk[q] + ((plus) && a == 1), ql[q] + ((plus) && a == 2))
    int *qj = (int *)malloc(nf * sizeof(int)), *qk = (int *)malloc(nf * sizeof(int)), *ql = (int *)malloc(nf * sizeof(int));
    for (int l = 1; l <= fm[2]; l++)
        for (int k = 1; k <= fm[1]; k++)
            for (int j = 1; j <= fm[0]; j++) {
                size_t q = (size_t)(j - 1) + (size_t)fm[0] * ((k - 1) + (size_t)fm[1] * (l - 1));
                qj[q] = a == 0 ? irup : j;
                qk[q] = a == 1 ? irup : k;
                ql[q] = a == 2 ? irup : l;
            }
    real *mrm = ralloc(nf), *mrp = ralloc(nf);
    for (size_t q = 0; q < nf; q++) {
        mrm[q] = AT(mr, qj[q], qk[q], ql[q]);
        mrp[q] = AT(mr, qj[q] + (a == 0), qk[q] + (a == 1), ql[q] + (a == 2));
    }
    /* trial traction for zero velocity and zero displacement */
    for (size_t q = 0; q < nf; q++) f1.p[q] = dt * dt * area.p[q] * (mrm[q] + mrp[q]);
    rinvert(f1.p, nf);
    for (int i = 1; i <= 3; i++)
        for (size_t q = 0; q < nf; q++) {
            real *T0 = t0 + (i - 1) * nf, *T1 = t1 + (i - 1) * nf, *T2 = t2 + (i - 1) * nf;
            T1[q] = T0[q] + f1.p[q] * dt *
                                (SIDE(vv, i, q, 1) - SIDE(vv, i, q, 0) + SIDE(w1, i, q, 1) * mrp[q] * dt -
                                 SIDE(w1, i, q, 0) * mrm[q] * dt);
            T2[q] = T1[q] + f1.p[q] * (SIDE(uu, i, q, 1) - SIDE(uu, i, q, 0));
        }
    /* shear and normal traction */
    for (size_t q = 0; q < nf; q++) tn.p[q] = t1[q] * nhat[q] + t1[nf + q] * nhat[nf + q] + t1[2 * nf + q] * nhat[2 * nf + q];
    for (int i = 0; i < 3; i++)
        for (size_t q = 0; q < nf; q++) t3[i * nf + q] = t1[i * nf + q] - tn.p[q] * nhat[i * nf + q];
    vnorm3(ts, t3);
    if (it > 1) {
        for (size_t q = 0; q < nf; q++) {
            tn.p[q] = t2[q] * nhat[q] + t2[nf + q] * nhat[nf + q] + t2[2 * nf + q] * nhat[2 * nf + q];
            if (faultopening == 1) tn.p[q] = fminr(0, tn.p[q]);
        }
        for (int i = 1; i <= 3; i++)
            for (size_t q = 0; q < nf; q++) t2[(i - 1) * nf + q] = SIDE(vv, i, q, 1) - SIDE(vv, i, q, 0);
        for (size_t q = 0; q < nf; q++) f2.p[q] = t2[q] * t2[q] + t2[nf + q] * t2[nf + q] + t2[2 * nf + q] * t2[2 * nf + q];
        /* slip-weakening friction law */
        for (size_t q = 0; q < nf; q++) {
            f1.p[q] = mud.p[q];
            if (sl.p[q] < dc.p[q]) f1.p[q] = f1.p[q] + (R(1.0) - sl.p[q] / dc.p[q]) * (mus.p[q] - mud.p[q]);
            f1.p[q] = -(fminr(0, tn.p[q]) * f1.p[q]) + co.p[q];
        }
        /* nucleation */
        if (rcrit > 0 && vrup > 0) {
            for (size_t q = 0; q < nf; q++) {
                f2.p[q] = 1;
                if (trelax > 0) f2.p[q] = fminr((tm - rhypo.p[q] / vrup) / trelax, R(1.0));
                f2.p[q] = (R(1.0) - f2.p[q]) * ts.p[q] + f2.p[q] * (-(tn.p[q] * mud.p[q]) + co.p[q]);
                if (rhypo.p[q] < fminr(rcrit, tm * vrup) && f2.p[q] < f1.p[q]) f1.p[q] = f2.p[q];
            }
        }
        /* shear traction bounded by friction */
        for (size_t q = 0; q < nf; q++) {
            f2.p[q] = 1;
            if (ts.p[q] > f1.p[q]) f2.p[q] = f1.p[q] / ts.p[q];
        }
        for (int i = 0; i < 3; i++)
            for (size_t q = 0; q < nf; q++) t3[i * nf + q] = f2.p[q] * t3[i * nf + q];
        for (size_t q = 0; q < nf; q++) ts.p[q] = fminr(ts.p[q], f1.p[q]);
        for (int i = 0; i < 3; i++)
            for (size_t q = 0; q < nf; q++) t1[i * nf + q] = t3[i * nf + q] + tn.p[q] * nhat[i * nf + q];
    }
    /* update acceleration */
    for (int i = 1; i <= 3; i++)
        for (size_t q = 0; q < nf; q++) {
            f2.p[q] = area.p[q] * (t1[(i - 1) * nf + q] - t0[(i - 1) * nf + q]);
            SIDE(w1, i, q, 0) = SIDE(w1, i, q, 0) + f2.p[q];
            SIDE(w1, i, q, 1) = SIDE(w1, i, q, 1) - f2.p[q];
        }
    vector_bc(w1, nm, bc1, bc2, i1bc, i2bc);
    field_io(">", "tx", fcomp(t1, 1));
    field_io(">", "ty", fcomp(t1, 2));
    field_io(">", "tz", fcomp(t1, 3));
    field_io(">", "tsx", fcomp(t3, 1));
    field_io(">", "tsy", fcomp(t3, 2));
    field_io(">", "tsz", fcomp(t3, 3));
    field_io(">", "tsm", ts);
    field_io(">", "tnm", tn);
    field_io(">", "fr", f1);
    set_halo(ts, -1, i1core, i2core);
    tsmax = maxval_(ts);
    set_halo(tn, HUGE_R, i1core, i2core);
    tnmin = minval_(tn);
    set_halo(tn, -HUGE_R, i1core, i2core);
    tnmax = maxval_(tn);
    set_halo(tn, 0, i1core, i2core);
    /* friction + fracture energy */
    for (int i = 1; i <= 3; i++)
        for (size_t q = 0; q < nf; q++) t2[(i - 1) * nf + q] = SIDE(vv, i, q, 1) - SIDE(vv, i, q, 0);
    for (size_t q = 0; q < nf; q++)
        f2.p[q] = (t1[q] * t2[q] + t1[nf + q] * t2[nf + q] + t1[2 * nf + q] * t2[2 * nf + q]) * area.p[q];
    set_halo(f2, 0, i1core, i2core);
    efric = efric + dt * sum_(f2);
    /* strain energy */
    for (int i = 1; i <= 3; i++)
        for (size_t q = 0; q < nf; q++) t2[(i - 1) * nf + q] = SIDE(uu, i, q, 1) - SIDE(uu, i, q, 0);
    for (size_t q = 0; q < nf; q++)
        f2.p[q] = ((t0[q] + t1[q]) * t2[q] + (t0[nf + q] + t1[nf + q]) * t2[nf + q] +
                   (t0[2 * nf + q] + t1[2 * nf + q]) * t2[2 * nf + q]) *
                  area.p[q];
    set_halo(f2, 0, i1core, i2core);
    estrain = R(0.5) * sum_(f2);
    /* moment */
    for (size_t q = 0; q < nf; q++)
        f2.p[q] = muf.p[q] * area.p[q] * SQRT(t2[q] * t2[q] + t2[nf + q] * t2[nf + q] + t2[2 * nf + q] * t2[2 * nf + q]);
    set_halo(f2, 0, i1core, i2core);
    moment = sum_(f2);
    /* slip acceleration */
    for (int i = 1; i <= 3; i++)
        for (size_t q = 0; q < nf; q++)
            t2[(i - 1) * nf + q] = SIDE(w1, i, q, 1) * mrp[q] - SIDE(w1, i, q, 0) * mrm[q];
    vnorm3(f2, t2);
    field_io(">", "sax", fcomp(t2, 1));
    field_io(">", "say", fcomp(t2, 2));
    field_io(">", "saz", fcomp(t2, 3));
    field_io(">", "sam", f2);
    set_halo(f2, -1, i1core, i2core);
    samax = maxval_(f2);
    free(mrm);
    free(mrp);
    free(qj);
    free(qk);
    free(ql);
#undef SIDE
}

/* ------------------------------------------------------------------------ */
/* step_accel: sord/acceleration.f90:6-265                                  */
/* ------------------------------------------------------------------------ */
static void step_accel(void) {
    int i1[3], i2[3], one[3] = {1, 1, 1};
    set_halo(s1, 0, i1node, i2node);
    for (int ic = 1; ic <= 3; ic++)
        for (int iid = 1; iid <= 3; iid++) {
            int id = fmodulo(ic + iid - 2, 3) + 1;
            for (int a = 0; a < 3; a++) i1[a] = i1node[a], i2[a] = i2node[a];
            if (ic == id)
                diff_cn(s1, w1, ic, id, i1, i2, xx);
            else
                diff_cn(s1, w2, 6 - ic - id, id, i1, i2, xx);
            /* pml region: p'_ij + d_j*p_ij = w_ij,j; f_i = sum_j(p_ij') */
            int a = id - 1;
            {
                A3 pp = slab(a == 0 ? p1 : a == 1 ? p2 : p3, a, ic, 0);
                int lo[3] = {i1[0], i1[1], i1[2]}, hi[3] = {i2[0], i2[1], i2[2]};
                hi[a] = imin(i2[a], i1pml[a]);
                for (int l = lo[2]; l <= hi[2]; l++)
                    for (int k = lo[1]; k <= hi[1]; k++)
                        for (int j = lo[0]; j <= hi[0]; j++) {
                            int c = a == 0 ? j : a == 1 ? k : l;
                            int i = c - i1[a] + 1, p = c + nnoff[a];
                            int pj = a == 0 ? i : j, pk = a == 1 ? i : k, pl = a == 2 ? i : l;
                            AT(s1, j, k, l) = dn2[p - 1] * AT(s1, j, k, l) + dn1[p - 1] * AT(pp, pj, pk, pl);
                            AT(pp, pj, pk, pl) = AT(pp, pj, pk, pl) + dt * AT(s1, j, k, l);
                        }
                pp = slab(a == 0 ? p4 : a == 1 ? p5 : p6, a + 3, ic, 0);
                for (int b = 0; b < 3; b++) lo[b] = i1[b], hi[b] = i2[b];
                lo[a] = imax(i1[a], i2pml[a]);
                for (int l = lo[2]; l <= hi[2]; l++)
                    for (int k = lo[1]; k <= hi[1]; k++)
                        for (int j = lo[0]; j <= hi[0]; j++) {
                            int c = a == 0 ? j : a == 1 ? k : l;
                            int i = i2[a] - c + 1, p = nn[a] - c - nnoff[a] + 1;
                            int pj = a == 0 ? i : j, pk = a == 1 ? i : k, pl = a == 2 ? i : l;
                            AT(s1, j, k, l) = dn2[p - 1] * AT(s1, j, k, l) + dn1[p - 1] * AT(pp, pj, pk, pl);
                            AT(pp, pj, pk, pl) = AT(pp, pj, pk, pl) + dt * AT(s1, j, k, l);
                        }
            }
            /* add contribution to force vector */
            real *W = w1 + (ic - 1) * vol_n;
            if (ic == id)
                memcpy(W, s1.p, vol_n * sizeof(real));
            else {
#pragma omp parallel for schedule(static)
                for (size_t q = 0; q < vol_n; q++) W[q] = W[q] + s1.p[q];
            }
        }
    /* hourglass control, only viscous in pml */
    if (hourglass[0] > 0 || hourglass[1] > 0) {
        set_halo(s1, 0, i1cell, i2cell);
        set_halo(s2, 0, i1node, i2node);
        for (int i = 0; i < 3; i++) {
            real *W = w2 + i * vol_n, *U = uu + i * vol_n, *V = vv + i * vol_n;
#pragma omp parallel for schedule(static)
            for (size_t q = 0; q < vol_n; q++) W[q] = hourglass[0] * U[q] + dt * hourglass[1] * V[q];
        }
        for (int iq = 1; iq <= 4; iq++)
            for (int ic = 1; ic <= 3; ic++) {
                for (int a = 0; a < 3; a++) {
                    i1[a] = imax(i1pml[a], i1cell[a]);
                    i2[a] = imin(i2pml[a] - 1, i2cell[a]);
                }
                hourglass_nc(s1, w2, iq, ic, i1, i2);
#pragma omp parallel for schedule(static)
                for (size_t q = 0; q < vol_n; q++) s1.p[q] = s1.p[q] * yy.p[q];
                for (int a = 0; a < 3; a++) {
                    i1[a] = imax(i1pml[a] + 1, i1node[a]);
                    i2[a] = imin(i2pml[a] - 1, i2node[a]);
                }
                hourglass_cn(s2, s1, iq, i1, i2);
                if (hourglass[1] > 0 && npml > 0) {
                    for (int i = 0; i < 3; i++) {
                        for (int pass = 0; pass < 2; pass++) {
                            for (int a = 0; a < 3; a++) i1[a] = i1cell[a], i2[a] = i2cell[a];
                            if (pass == 0)
                                i2[i] = imin(i2[i], i1pml[i]);
                            else
                                i1[i] = imax(i1[i], i2pml[i] - 1);
                            hourglass_nc(s1, vv, iq, ic, i1, i2);
                            for (int l = i1[2]; l <= i2[2]; l++)
                                for (int k = i1[1]; k <= i2[1]; k++)
                                    for (int j = i1[0]; j <= i2[0]; j++)
                                        AT(s1, j, k, l) = dt * hourglass[1] * AT(yy, j, k, l) * AT(s1, j, k, l);
                        }
                    }
                    for (int i = 0; i < 3; i++) {
                        for (int a = 0; a < 3; a++) i1[a] = i1node[a], i2[a] = i2node[a];
                        i2[i] = imin(i2[i], i1pml[i]);
                        hourglass_cn(s2, s1, iq, i1, i2);
                        for (int a = 0; a < 3; a++) i1[a] = i1node[a], i2[a] = i2node[a];
                        i1[i] = imax(i1[i], i2pml[i]);
                        hourglass_cn(s2, s1, iq, i1, i2);
                    }
                }
                real *W = w1 + (ic - 1) * vol_n;
#pragma omp parallel for schedule(static)
                for (size_t q = 0; q < vol_n; q++) W[q] = W[q] - s2.p[q];
            }
    }
    if (!strcmp(source, "force")) finite_source();
    field_io("<", "fx", comp(w1, 1));
    field_io("<", "fy", comp(w1, 2));
    field_io("<", "fz", comp(w1, 3));
    vector_bc(w1, nm, bc1, bc2, i1bc, i2bc);
    step_rupture();
    timers[0] = clock_ms();
    timers[0] = clock_ms() - timers[0];
    field_io(">", "fx", comp(w1, 1));
    field_io(">", "fy", comp(w1, 2));
    field_io(">", "fz", comp(w1, 3));
    for (int i = 0; i < 3; i++) {
        real *W = w1 + i * vol_n;
#pragma omp parallel for schedule(static)
        for (size_t q = 0; q < vol_n; q++) W[q] = W[q] * mr.p[q];
    }
    field_io("<>", "ax", comp(w1, 1));
    field_io("<>", "ay", comp(w1, 2));
    field_io("<>", "az", comp(w1, 3));
    if (fmodulo(it, itstats) == 0) {
        vector_norm(s1, w1, i1core, i2core, one);
        set_halo(s1, -1, i1core, i2core);
        amax = maxloc_(s1, amaxloc);
    }
    field_io(">", "am2", s1);
}

/* ------------------------------------------------------------------------ */
/* stats: sord/statistics.f90:13-131                                        */
/* ------------------------------------------------------------------------ */
static void stats(void) {
    static int init = 1, j = 0, nmax;
    static real *maxl, *suml;
    static int32_t *prof;
    int rupt = strcmp(faultnormal, "no") != 0;
    timers[2] = clock_ms();
    if (init) {
        init = 0;
        nmax = rupt ? 12 : 4;
        maxl = ralloc((size_t)nmax * itio);
        suml = ralloc((size_t)3 * itio);
        prof = (int32_t *)calloc((size_t)4 * itio, sizeof(int32_t));
    }
    if (fmodulo(it, itstats) == 0) {
        j++;
        real *m = maxl + (size_t)nmax * (j - 1);
        m[0] = amax;
        m[1] = vmax;
        m[2] = umax;
        m[3] = wmax;
        real val = m[0];
        for (int q = 0; q < nmax * j; q++)
            if (maxl[q] > val || maxl[q] != maxl[q]) val = maxl[q];
        if (val != val || val > HUGE_R) die(1, "NaN/Inf!");
        if (rupt) {
            m[4] = samax;
            m[5] = svmax;
            m[6] = sumax;
            m[7] = slmax;
            m[8] = tsmax;
            m[9] = -tnmin;
            m[10] = tnmax;
            m[11] = tarrmax;
            real *s = suml + 3 * (size_t)(j - 1);
            s[0] = efric;
            s[1] = estrain;
            s[2] = moment;
        }
    }
    if (j > 0 && (fmodulo(it, itio) == 0 || it == nt)) {
        int m = nt / itstats, o = it / itstats - j;
        (void)m;
        static const char *names[12] = {"stats-amax.bin", "stats-vmax.bin", "stats-umax.bin", "stats-wmax.bin",
                                        "stats-samax.bin", "stats-svmax.bin", "stats-sumax.bin", "stats-slmax.bin",
                                        "stats-tsmax.bin", "stats-tnmin.bin", "stats-tnmax.bin", "stats-tarrmax.bin"};
        real *buf = ralloc(j);
        for (int q = 0; q < nmax; q++) {
            for (int r = 0; r < j; r++) {
                real v = maxl[(size_t)nmax * r + q];
                if (q < 4) v = SQRT(v);
                if (q == 9) v = -v;
                buf[r] = v;
            }
            rio_write(names[q], buf, 1, o, j);
        }
        if (rupt) {
            static const char *sn[3] = {"stats-efric.bin", "stats-estrain.bin", "stats-moment.bin"};
            for (int q = 0; q < 3; q++) {
                for (int r = 0; r < j; r++) buf[r] = suml[3 * (size_t)r + q];
                rio_write(sn[q], buf, 1, o, j);
            }
            for (int r = 0; r < j; r++) {
                real v = suml[3 * (size_t)r + 2];
                buf[r] = v > 0 ? ((real)log10((double)v) - R(9.05)) / R(1.5) : R(-999.0);
            }
            rio_write("stats-mw.bin", buf, 1, o, j);
        }
        free(buf);
        j = 0;
    }
    timers[3] = clock_ms();
    timers[2] = timers[3] - timers[2];
    int i = fmodulo(it - 1, itio) + 1;
    for (int q = 0; q < 4; q++) prof[4 * (i - 1) + q] = (int32_t)timers[q];
    if (i == itio || it == nt) {
        static const char *pn_[4] = {"prof-halo.bin", "prof-io.bin", "prof-stats.bin", "prof-step.bin"};
        for (int q = 0; q < 4; q++) {
            int32_t *b = (int32_t *)calloc(i, sizeof(int32_t));
            for (int r = 0; r < i; r++) b[r] = prof[4 * r + q];
            FILE *fp = fopen(pn_[q], it - i > 0 ? "r+b" : "wb");
            if (fp) {
                fseek(fp, (long)(it - i) * 4, SEEK_SET);
                fwrite(b, 4, i, fp);
                fclose(fp);
            }
            free(b);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* program sord: sord/sord.f90:24-47                                        */
/* ------------------------------------------------------------------------ */
int main(void) {
    clock_gettime(CLOCK_MONOTONIC, &clock0);
    read_parameters();
#ifdef _OPENMP
    if (nthread > 0) omp_set_num_threads(nthread);
#endif
    if (debug_ > 3) fprintf(stderr, "oracle: debug dumps are not restated\n");
    setup_dimensions();
    vol_n = (size_t)nm[0] * nm[1] * nm[2];
    xx = ralloc(3 * vol_n);
    vv = ralloc(3 * vol_n);
    uu = ralloc(3 * vol_n);
    w1 = ralloc(3 * vol_n);
    w2 = ralloc(3 * vol_n);
    vc = vol3(ralloc(vol_n));
    mr = vol3(ralloc(vol_n));
    lam = vol3(ralloc(vol_n));
    mu = vol3(ralloc(vol_n));
    gam = vol3(ralloc(vol_n));
    yy = vol3(ralloc(vol_n));
    s1 = vol3(ralloc(vol_n));
    s2 = vol3(ralloc(vol_n));
    for (int a = 0; a < 3; a++) fm[a] = nm[a];
    init_grid();
    init_material();
    init_pml();
    init_finite_source();
    init_rupture();
    resample_material();
    struct timespec ta, tb;
    clock_gettime(CLOCK_MONOTONIC, &ta);
    for (it = 1; it <= nt; it++) {
        clock_gettime(CLOCK_MONOTONIC, &clock0); /* per-step clock (sord/sord.f90:40) */
        timers[0] = timers[1] = timers[2] = timers[3] = 0;
        step_time();
        step_stress();
        step_accel();
        stats();
    }
    clock_gettime(CLOCK_MONOTONIC, &tb);
    double sec = (double)(tb.tv_sec - ta.tv_sec) + 1e-9 * (double)(tb.tv_nsec - ta.tv_nsec);
    int threads = 1;
#ifdef _OPENMP
    threads = omp_get_max_threads();
#endif
    printf("{\"loop_seconds\": %.6f, \"steps\": %d, \"nodes\": %ld, \"threads\": %d}\n", sec, nt,
           (long)nn[0] * nn[1] * nn[2], threads);
    return 0;
}
