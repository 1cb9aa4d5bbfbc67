/*
 * ifos_oracle.c -- CPU restatement of the IFOS3D hot path (TEST INFRASTRUCTURE).
 *
 * Only the test suite, the smoke check and the benchmark's CPU-baseline leg
 * use this file, and only as the checker.  See ifos_oracle.h for the layout.
 *
 * Every function restates one reference routine (cited file:line, paths
 * relative to the reference's src/).  Where the reference mixes float and
 * double (libm calls, double literals) the same promotions are written out
 * here, so the ORACLE_REAL=float build reproduces the reference's rounding
 * step for step.  Built with -O2 -ffp-contract=off and no -march flag: no
 * fused multiply-add, no excess precision, exactly like the reference's own
 * -O3 build on x86-64 (SSE2, FLT_EVAL_METHOD 0).
 */
#include "ifos_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define HALO_LO 2
#define HALO_HI 2

/* threads the OpenMP cell loops use (1 in a build without OpenMP) */
int og_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

size_t og_field_size(const og_grid *g)
{
    return (size_t)(g->ny + 5) * (size_t)(g->nx + 5) * (size_t)(g->nz + 5);
}

size_t og_idx(const og_grid *g, int j, int i, int k)
{
    return ((size_t)(j + HALO_LO) * (size_t)(g->nx + 5) + (size_t)(i + HALO_LO)) *
               (size_t)(g->nz + 5) +
           (size_t)(k + HALO_LO);
}

/* strides of the canonical layout */
#define SJ(g) ((ptrdiff_t)((g)->nx + 5) * (ptrdiff_t)((g)->nz + 5))
#define SI(g) ((ptrdiff_t)((g)->nz + 5))

/* ------------------------------------------------------------------ */
/* FD coefficients, update_v_ssg.c:98-100 (and every 4th-order kernel) */
static void fd4_coeffs(int fdcoeff, oreal *b1, oreal *b2)
{
    *b1 = (oreal)(9.0 / 8.0);
    *b2 = (oreal)(-1.0 / 24.0);
    if (fdcoeff == 2) {
        *b1 = (oreal)1.1382;
        *b2 = (oreal)(-0.046414);
    }
}

/* CPML_ini_elastic.c:26-71 with POS = {0,0,0}, NPROC* = 1 */
void og_interior_box(const og_grid *g, int box[6])
{
    box[0] = g->fw + 1;
    box[1] = g->nx - g->fw;
    box[2] = g->free_surf ? 1 : g->fw + 1;
    box[3] = g->ny - g->fw;
    box[4] = g->fw + 1;
    box[5] = g->nz - g->fw;
}

/* ------------------------------------------------------------------ */
/* CPML_coeff.c:87-324.  The reference tests |d| with the integer abs()
 * (the float argument is truncated), so "abs(d) > 1e-6" means |d| >= 1.
 * The y loop's "else" branch writes a_x[i] = 0 instead of a_y[i]
 * (CPML_coeff.c:216); that is reproduced in og_cpml_coeff. */
static int int_abs_positive(oreal d)
{
    int t = (int)d;
    return abs(t) > 1.0e-6;
}

static void cpml_one_axis(const og_grid *g, double dh, float vppml,
                          float fpml, oreal *K, oreal *alpha, oreal *a,
                          oreal *b, oreal *K_h, oreal *alpha_h, oreal *a_h,
                          oreal *b_h, oreal *d)
{
    const int fw = g->fw;
    const oreal npower = 3;
    const oreal k_max_PML = 2.0;
    const oreal alpha_max_PML = (oreal)(2.0 * 3.141592653589793 * (fpml / 2.0));
    const oreal Rcoef = (oreal)0.0008;
    const oreal ca = (oreal)0.25, cb = (oreal)0.75, cc = (oreal)0.0;
    const oreal dhr = (oreal)dh;
    oreal *d_h = (oreal *)calloc((size_t)(2 * fw + 2), sizeof(oreal));
    oreal d0, pos, pn;
    int i;

    d0 = (oreal)(-(npower + 1) * vppml * log(Rcoef) / (2.0 * fw * dhr));

    for (i = 1; i <= fw + 1; i++) {
        K[i] = 1.0;
        pos = (oreal)((fw - i + 1) * dhr);
        pn = pos / (oreal)(fw * dhr);
        d[i] = (oreal)(d0 * (ca * pn + cb * pow(pn, npower) + cc * pow(pn, 4)));
        K[i] = (oreal)(1.0 + (k_max_PML - 1.0) * pow(pn, npower));
        alpha[i] = (oreal)(alpha_max_PML * (1.0 - pn));
        b[i] = (oreal)exp(-(d[i] / K[i] + alpha[i]) * (oreal)g->dt);
        if (int_abs_positive(d[i]))
            a[i] = (oreal)(d[i] * (b[i] - 1.0) / (K[i] * (d[i] + K[i] * alpha[i])));
        if (i <= fw) {
            int l = 2 * fw - i + 1;
            pos = (oreal)((fw - i + 0.5) * dhr);
            pn = pos / (oreal)(fw * dhr);
            K_h[i] = 1.0;
            d_h[i] = (oreal)(d0 * (ca * pn + cb * pow(pn, npower) + cc * pow(pn, 4)));
            K_h[i] = (oreal)(1.0 + (k_max_PML - 1.0) * pow(pn, npower));
            alpha_h[i] = (oreal)(alpha_max_PML * (1.0 - pn));
            b_h[i] = (oreal)exp(-(d_h[i] / K_h[i] + alpha_h[i]) * (oreal)g->dt);
            if (int_abs_positive(d_h[i]))
                a_h[i] = (oreal)(d_h[i] * (b_h[i] - 1.0) /
                                 (K_h[i] * (d_h[i] + K_h[i] * alpha_h[i])));
            /* right side mirrors the left; integer-grid profiles land one
             * cell further out than the half-grid ones (CPML_coeff.c:163-186) */
            if (i > 1) {
                K[l + 1] = K[i];
                b[l + 1] = b[i];
                if (int_abs_positive(d[i])) a[l + 1] = a[i];
            }
            K_h[l] = K_h[i];
            b_h[l] = b_h[i];
            if (int_abs_positive(d[i])) a_h[l] = a_h[i];
        }
    }
    free(d_h);
}

void og_cpml_coeff(const og_grid *g, float vppml, float fpml, oreal *prof[18],
                   oreal *alpha[6])
{
    int ax, i;
    const float dh[3] = {g->dx, g->dy, g->dz};
    oreal *d = (oreal *)calloc((size_t)(2 * g->fw + 2), sizeof(oreal));
    for (ax = 0; ax < 3; ax++) {
        oreal **p = prof + 6 * ax;
        int n;
        for (n = 0; n < 6; n++) memset(p[n], 0, sizeof(oreal) * (size_t)(2 * g->fw + 1));
        memset(alpha[2 * ax], 0, sizeof(oreal) * (size_t)(2 * g->fw + 1));
        memset(alpha[2 * ax + 1], 0, sizeof(oreal) * (size_t)(2 * g->fw + 1));
        memset(d, 0, sizeof(oreal) * (size_t)(2 * g->fw + 2));
        cpml_one_axis(g, (double)dh[ax], vppml, fpml, p[0], alpha[2 * ax], p[1], p[2],
                      p[3], alpha[2 * ax + 1], p[4], p[5], d);
        /* the y pass clears a_x where its own |d_y| < 1 (CPML_coeff.c:215-216) */
        if (ax == 1)
            for (i = 1; i <= g->fw + 1; i++)
                if (!int_abs_positive(d[i])) prof[1][i] = 0.0;
    }
    free(d);
}

/* ------------------------------------------------------------------ */
/* matcopy.c:31-307 with every neighbour MPI_PROC_NULL: the send buffer is
 * received back unchanged, so each halo face takes the opposite interior
 * face.  Done in the reference's order (y faces, then x, then z) over the
 * reference's index ranges, which leaves a fully periodic halo. */
void og_matcopy_wrap(const og_grid *g, oreal *m)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    int i, j, k;
    for (i = 0; i <= nx + 1; i++)
        for (k = 0; k <= nz + 1; k++) {
            oreal top = m[og_idx(g, 1, i, k)], bot = m[og_idx(g, ny, i, k)];
            m[og_idx(g, ny + 1, i, k)] = top;
            m[og_idx(g, 0, i, k)] = bot;
        }
    for (j = 0; j <= ny + 1; j++)
        for (k = 0; k <= nz + 1; k++) {
            oreal lef = m[og_idx(g, j, 1, k)], rig = m[og_idx(g, j, nx, k)];
            m[og_idx(g, j, nx + 1, k)] = lef;
            m[og_idx(g, j, 0, k)] = rig;
        }
    for (i = 0; i <= nx + 1; i++)
        for (j = 0; j <= ny + 1; j++) {
            oreal fro = m[og_idx(g, j, i, 1)], bac = m[og_idx(g, j, i, nz)];
            m[og_idx(g, j, i, nz + 1)] = fro;
            m[og_idx(g, j, i, 0)] = bac;
        }
}

/* av_mat.c:28-80 */
void og_av_mat(const og_grid *g, const oreal *rho, const oreal *u,
               const oreal *taus, oreal *uipjp, oreal *ujpkp, oreal *uipkp,
               oreal *tausipjp, oreal *tausjpkp, oreal *tausipkp, oreal *rjp,
               oreal *rkp, oreal *rip)
{
    const ptrdiff_t sj = SJ(g), si = SI(g);
    int i, j, k;
    for (j = 1; j <= g->ny; j++)
        for (i = 1; i <= g->nx; i++)
            for (k = 1; k <= g->nz; k++) {
                size_t c = og_idx(g, j, i, k);
                const oreal *uc = u + c;
                uipjp[c] = (oreal)(4.0 / ((1.0 / uc[0]) + (1.0 / uc[si]) +
                                          (1.0 / uc[sj + si]) + (1.0 / uc[sj])));
                ujpkp[c] = (oreal)(4.0 / ((1.0 / uc[0]) + (1.0 / uc[1]) +
                                          (1.0 / uc[sj + 1]) + (1.0 / uc[sj])));
                uipkp[c] = (oreal)(4.0 / ((1.0 / uc[0]) + (1.0 / uc[1]) +
                                          (1.0 / uc[si + 1]) + (1.0 / uc[si])));
                if (taus) {
                    const oreal *tc = taus + c;
                    tausipjp[c] = (oreal)(0.25 * (tc[0] + tc[si] + tc[sj + si] + tc[sj]));
                    tausjpkp[c] = (oreal)(0.25 * (tc[0] + tc[sj] + tc[sj + 1] + tc[1]));
                    tausipkp[c] = (oreal)(0.25 * (tc[0] + tc[si] + tc[si + 1] + tc[1]));
                }
                rjp[c] = (oreal)(0.5 * (rho[c] + rho[c + sj]));
                rkp[c] = (oreal)(0.5 * (rho[c] + rho[c + 1]));
                rip[c] = (oreal)(0.5 * (rho[c] + rho[c + si]));
            }
}

/* ------------------------------------------------------------------ */
/* wavelet.c:47-90 (shapes 1, 2, 4, 5; shape 3 reads a file and is not
 * part of this path) */
void og_wavelet(const og_grid *g, const og_source *src, int nsrc, int shape,
                float *signals)
{
    const double PI_ = 3.141592653589793;
    int nt, s;
    for (nt = 1; nt <= g->nt; nt++) {
        float t = (float)nt * g->dt;
        for (s = 0; s < nsrc; s++) {
            float tshift = src[s].tshift, fc = src[s].fc, a = src[s].amp;
            float ts = (float)(1.0 / fc), tau, amp = 0.0f;
            switch (shape) {
            case 1:
                tau = (float)(PI_ * (t - 1.5 * ts - tshift) / (ts));
                amp = (float)((1.0 - 2.0 * tau * tau) * exp(-tau * tau));
                break;
            case 2:
                if ((t < tshift) || (t > (tshift + ts)))
                    amp = 0.0f;
                else
                    amp = (float)(sin(2.0 * PI_ * (t - tshift) * fc) -
                                  0.5 * sin(4.0 * PI_ * (t - tshift) * fc));
                break;
            case 4:
                if ((t < tshift) || (t > (tshift + ts)))
                    amp = 0.0f;
                else
                    amp = (float)((0.75 * PI_ / ts) * pow(sin(PI_ * (t - tshift) / ts), 3.0));
                break;
            case 5:
                amp = (t < tshift) ? 0.0f : 1.0f;
                break;
            default:
                amp = 0.0f;
            }
            signals[(size_t)s * (size_t)(g->nt + 2) + (size_t)nt] = amp * a;
        }
    }
}

/* ------------------------------------------------------------------ */
/* 4th-order staggered differences.  Forward difference at c: f(c+1)-f(c),
 * backward: f(c)-f(c-1); the stencil's long arm is two cells out. */
#define DFWD(f, c, s) (b1 * ((f)[(c) + (s)] - (f)[(c)]) + b2 * ((f)[(c) + 2 * (s)] - (f)[(c) - (s)]))
#define DBWD(f, c, s) (b1 * ((f)[(c)] - (f)[(c) - (s)]) + b2 * ((f)[(c) + (s)] - (f)[(c) - 2 * (s)]))

/* the 9 stress derivatives of the velocity update, update_v_ssg.c:106-123 */
typedef struct { oreal sxx_x, sxy_y, sxz_z, syy_y, sxy_x, syz_z, szz_z, sxz_x, syz_y; } sdiff;

static inline void stress_diffs(const og_grid *g, size_t c, const oreal *sxx,
                                const oreal *syy, const oreal *szz,
                                const oreal *sxy, const oreal *syz,
                                const oreal *sxz, oreal b1, oreal b2, sdiff *d)
{
    const ptrdiff_t sj = SJ(g), si = SI(g);
    const oreal dx = (oreal)g->dt / (oreal)g->dx, dy = (oreal)g->dt / (oreal)g->dy,
                dz = (oreal)g->dt / (oreal)g->dz;
    d->sxx_x = dx * DFWD(sxx, c, si);
    d->sxy_y = dy * DBWD(sxy, c, sj);
    d->sxz_z = dz * DBWD(sxz, c, 1);
    d->syy_y = dy * DFWD(syy, c, sj);
    d->sxy_x = dx * DBWD(sxy, c, si);
    d->syz_z = dz * DBWD(syz, c, 1);
    d->szz_z = dz * DFWD(szz, c, 1);
    d->sxz_x = dx * DBWD(sxz, c, si);
    d->syz_y = dy * DBWD(syz, c, sj);
}

static void inject_forces(const og_grid *g, int nt, oreal *vx, oreal *vy,
                          oreal *vz, const oreal *rjp, const oreal *rkp,
                          const oreal *rip, const og_source *srcs,
                          const float *sig, int nsrc)
{
    /* update_v_ssg.c:474-510: body forces.  Types 3 and 4 pair vz with rjp
     * and vy with rkp, and the density is multiplied, as the reference does. */
    const double PI_ = 3.141592653589793;
    int l;
    for (l = 0; l < nsrc; l++) {
        size_t c = og_idx(g, srcs[l].j, srcs[l].i, srcs[l].k);
        float amp = g->dt * sig[(size_t)l * (size_t)(g->nt + 2) + (size_t)nt] /
                    (g->dx * g->dy * g->dz);
        switch (srcs[l].type) {
        case 2: vx[c] += amp * rip[c]; break;
        case 3: vz[c] += amp * rjp[c]; break;
        case 4: vy[c] += amp * rkp[c]; break;
        case 5: {
            /* ALPHA = BETA = 0 on this path (globals not exposed) */
            float alpha_rad = (float)(0.0f * PI_ / 180), beta_rad = (float)(0.0f * PI_ / 180);
            vx[c] = (oreal)(vx[c] + cos(beta_rad) * cos(alpha_rad) * amp);
            vy[c] = (oreal)(vy[c] + sin(alpha_rad) * amp);
            vz[c] = (oreal)(vz[c] + sin(beta_rad) * cos(alpha_rad) * amp);
        } break;
        default: break;
        }
    }
}

static void inject_adjoint(const og_grid *g, int nt, oreal *vx, oreal *vy,
                           oreal *vz, const og_source *srcs, const float *sx,
                           const float *sy, const float *sz, int nsrc)
{
    /* update_v_ssg.c:512-522 and update_v_ssg_CPML.c:384-393 (raw add) */
    int l;
    for (l = 0; l < nsrc; l++) {
        size_t c = og_idx(g, srcs[l].j, srcs[l].i, srcs[l].k);
        size_t s = (size_t)l * (size_t)(g->nt + 2) + (size_t)nt;
        vx[c] += sx[s];
        vz[c] += sz[s];
        vy[c] += sy[s];
    }
}

void og_update_v(const og_grid *g, const int box[6], int nt, oreal *vx,
                 oreal *vy, oreal *vz, const oreal *sxx, const oreal *syy,
                 const oreal *szz, const oreal *sxy, const oreal *syz,
                 const oreal *sxz, const oreal *rjp, const oreal *rkp,
                 const oreal *rip, const og_source *srcs, const float *sig_x,
                 const float *sig_y, const float *sig_z, int nsrc, int back)
{
    oreal b1, b2;
    int i, j, k;
    sdiff d;
    fd4_coeffs(g->fdcoeff, &b1, &b2);
    #pragma omp parallel for private(i, k, d) schedule(static)
    for (j = box[2]; j <= box[3]; j++)
        for (i = box[0]; i <= box[1]; i++)
            for (k = box[4]; k <= box[5]; k++) {
                size_t c = og_idx(g, j, i, k);
                stress_diffs(g, c, sxx, syy, szz, sxy, syz, sxz, b1, b2, &d);
                vx[c] += (d.sxx_x + d.sxy_y + d.sxz_z) / rip[c];
                vy[c] += (d.syy_y + d.sxy_x + d.syz_z) / rjp[c];
                vz[c] += (d.szz_z + d.sxz_x + d.syz_y) / rkp[c];
            }
    if (back == 0)
        inject_forces(g, nt, vx, vy, vz, rjp, rkp, rip, srcs, sig_x, nsrc);
    else
        inject_adjoint(g, nt, vx, vy, vz, srcs, sig_x, sig_y, sig_z, nsrc);
}

/* slab bookkeeping shared by both CPML updates (update_v_ssg_CPML.c:70-375,
 * update_s_ssg_CPML_elastic.c:59-480).  A cell outside the interior box is
 * in the x slabs if i <= fw or i >= nx-fw+1 (all j, k), otherwise in the y
 * slabs if j is (all k), otherwise in the z slabs.  Inside a slab cell, the
 * psi recursion is applied for every axis along which the cell lies in a
 * damping layer; the slab index h maps [1..fw] and [n-fw+1..n] onto
 * [1..2fw]. */
static inline int slab_h(int idx, int n, int fw, int lo_active)
{
    if (lo_active && idx <= fw) return idx;
    if (idx >= n - fw + 1) return idx - (n - 2 * fw);
    return 0;
}

static inline size_t psi_x_at(const og_grid *g, int j, int h, int k)
{
    return ((size_t)(j - 1) * (size_t)(2 * g->fw) + (size_t)(h - 1)) * (size_t)g->nz + (size_t)(k - 1);
}
static inline size_t psi_y_at(const og_grid *g, int h, int i, int k)
{
    return ((size_t)(h - 1) * (size_t)g->nx + (size_t)(i - 1)) * (size_t)g->nz + (size_t)(k - 1);
}
static inline size_t psi_z_at(const og_grid *g, int j, int i, int h)
{
    return ((size_t)(j - 1) * (size_t)g->nx + (size_t)(i - 1)) * (size_t)(2 * g->fw) + (size_t)(h - 1);
}

static inline int in_box(const int box[6], int j, int i, int k)
{
    return i >= box[0] && i <= box[1] && j >= box[2] && j <= box[3] &&
           k >= box[4] && k <= box[5];
}

/* profile accessors: prof[6*axis + {0:K 1:a 2:b 3:K_half 4:a_half 5:b_half}] */
#define PK(ax, h) prof[6 * (ax) + 0][h]
#define PA(ax, h) prof[6 * (ax) + 1][h]
#define PB(ax, h) prof[6 * (ax) + 2][h]
#define PKH(ax, h) prof[6 * (ax) + 3][h]
#define PAH(ax, h) prof[6 * (ax) + 4][h]
#define PBH(ax, h) prof[6 * (ax) + 5][h]

/* psi <- b psi + a d ; d <- d / K + psi   (update_v_ssg_CPML.c:85-90) */
#define CPML_APPLY(psi, at, d, K, a, b) \
    do {                                 \
        (psi)[at] = (b) * (psi)[at] + (a) * (d); \
        (d) = (d) / (K) + (psi)[at];     \
    } while (0)

void og_update_v_cpml(const og_grid *g, const int box[6], int nt, oreal *vx,
                      oreal *vy, oreal *vz, const oreal *sxx,
                      const oreal *syy, const oreal *szz, const oreal *sxy,
                      const oreal *syz, const oreal *sxz, const oreal *rjp,
                      const oreal *rkp, const oreal *rip,
                      const og_source *srcs, const float *sig_x,
                      const float *sig_y, const float *sig_z, int nsrc,
                      int back, oreal *const prof[18], oreal *psi[9])
{
    oreal b1, b2;
    int i, j, k;
    const int fw = g->fw;
    const int ytop = !g->free_surf;
    sdiff d;
    oreal *p_sxx_x = psi[0], *p_sxy_x = psi[1], *p_sxz_x = psi[2];
    oreal *p_sxy_y = psi[3], *p_syy_y = psi[4], *p_syz_y = psi[5];
    oreal *p_sxz_z = psi[6], *p_syz_z = psi[7], *p_szz_z = psi[8];
    fd4_coeffs(g->fdcoeff, &b1, &b2);
    #pragma omp parallel for private(i, k, d) schedule(static)
    for (j = 1; j <= g->ny; j++)
        for (i = 1; i <= g->nx; i++)
            for (k = 1; k <= g->nz; k++) {
                int hx, hy, hz;
                size_t c;
                if (in_box(box, j, i, k)) continue;
                /* cells of the top y layer with a free surface belong to
                 * the interior update (yb0 = 1) */
                c = og_idx(g, j, i, k);
                stress_diffs(g, c, sxx, syy, szz, sxy, syz, sxz, b1, b2, &d);
                hx = slab_h(i, g->nx, fw, 1);
                hy = slab_h(j, g->ny, fw, ytop);
                hz = slab_h(k, g->nz, fw, 1);
                if (hx) {
                    CPML_APPLY(p_sxx_x, psi_x_at(g, j, hx, k), d.sxx_x, PKH(0, hx), PAH(0, hx), PBH(0, hx));
                    CPML_APPLY(p_sxy_x, psi_x_at(g, j, hx, k), d.sxy_x, PK(0, hx), PA(0, hx), PB(0, hx));
                    CPML_APPLY(p_sxz_x, psi_x_at(g, j, hx, k), d.sxz_x, PK(0, hx), PA(0, hx), PB(0, hx));
                }
                if (hy) {
                    CPML_APPLY(p_sxy_y, psi_y_at(g, hy, i, k), d.sxy_y, PK(1, hy), PA(1, hy), PB(1, hy));
                    CPML_APPLY(p_syy_y, psi_y_at(g, hy, i, k), d.syy_y, PKH(1, hy), PAH(1, hy), PBH(1, hy));
                    CPML_APPLY(p_syz_y, psi_y_at(g, hy, i, k), d.syz_y, PK(1, hy), PA(1, hy), PB(1, hy));
                }
                if (hz) {
                    CPML_APPLY(p_sxz_z, psi_z_at(g, j, i, hz), d.sxz_z, PK(2, hz), PA(2, hz), PB(2, hz));
                    CPML_APPLY(p_syz_z, psi_z_at(g, j, i, hz), d.syz_z, PK(2, hz), PA(2, hz), PB(2, hz));
                    CPML_APPLY(p_szz_z, psi_z_at(g, j, i, hz), d.szz_z, PKH(2, hz), PAH(2, hz), PBH(2, hz));
                }
                vx[c] += (d.sxx_x + d.sxy_y + d.sxz_z) / rip[c];
                vy[c] += (d.syy_y + d.sxy_x + d.syz_z) / rjp[c];
                vz[c] += (d.szz_z + d.sxz_x + d.syz_y) / rkp[c];
            }
    if (back == 1)
        inject_adjoint(g, nt, vx, vy, vz, srcs, sig_x, sig_y, sig_z, nsrc);
}

/* ------------------------------------------------------------------ */
/* the 9 velocity derivatives of the stress update,
 * update_s_ssg_elastic.c:106-114 */
typedef struct { oreal vxx, vxy, vxz, vyx, vyy, vyz, vzx, vzy, vzz; } vdiff;

static inline void velocity_diffs(const og_grid *g, size_t c, const oreal *vx,
                                  const oreal *vy, const oreal *vz, oreal b1,
                                  oreal b2, vdiff *d)
{
    const ptrdiff_t sj = SJ(g), si = SI(g);
    const oreal DX = g->dx, DY = g->dy, DZ = g->dz;
    d->vxx = DBWD(vx, c, si) / DX;
    d->vxy = DFWD(vx, c, sj) / DY;
    d->vxz = DFWD(vx, c, 1) / DZ;
    d->vyx = DFWD(vy, c, si) / DX;
    d->vyy = DBWD(vy, c, sj) / DY;
    d->vyz = DFWD(vy, c, 1) / DZ;
    d->vzx = DFWD(vz, c, si) / DX;
    d->vzy = DFWD(vz, c, sj) / DY;
    d->vzz = DBWD(vz, c, 1) / DZ;
}

/* update_s_ssg_elastic.c:117-138 */
static inline void stress_apply(const og_grid *g, size_t c, const vdiff *d,
                                oreal *sxx, oreal *syy, oreal *szz,
                                oreal *sxy, oreal *syz, oreal *sxz,
                                const oreal *pi, const oreal *u,
                                const oreal *uipjp, const oreal *ujpkp,
                                const oreal *uipkp)
{
    const oreal DT = g->dt;
    oreal fipjp = uipjp[c] * DT, fjpkp = ujpkp[c] * DT, fipkp = uipkp[c] * DT;
    oreal gg = pi[c];
    oreal f = (oreal)(2.0 * u[c]);
    oreal vxyyx = d->vxy + d->vyx, vyzzy = d->vyz + d->vzy, vxzzx = d->vxz + d->vzx;
    oreal vxxyyzz = d->vxx + d->vyy + d->vzz;
    oreal vyyzz = d->vyy + d->vzz, vxxzz = d->vxx + d->vzz, vxxyy = d->vxx + d->vyy;
    sxy[c] += (fipjp * vxyyx);
    syz[c] += (fjpkp * vyzzy);
    sxz[c] += (fipkp * vxzzx);
    sxx[c] += DT * ((gg * vxxyyzz) - (f * vyyzz));
    syy[c] += DT * ((gg * vxxyyzz) - (f * vxxzz));
    szz[c] += DT * ((gg * vxxyyzz) - (f * vxxyy));
}

void og_update_s_elastic(const og_grid *g, const int box[6], const oreal *vx,
                         const oreal *vy, const oreal *vz, oreal *sxx,
                         oreal *syy, oreal *szz, oreal *sxy, oreal *syz,
                         oreal *sxz, const oreal *pi, const oreal *u,
                         const oreal *uipjp, const oreal *ujpkp,
                         const oreal *uipkp)
{
    oreal b1, b2;
    int i, j, k;
    vdiff d;
    fd4_coeffs(g->fdcoeff, &b1, &b2);
    #pragma omp parallel for private(i, k, d) schedule(static)
    for (j = box[2]; j <= box[3]; j++)
        for (i = box[0]; i <= box[1]; i++)
            for (k = box[4]; k <= box[5]; k++) {
                size_t c = og_idx(g, j, i, k);
                velocity_diffs(g, c, vx, vy, vz, b1, b2, &d);
                stress_apply(g, c, &d, sxx, syy, szz, sxy, syz, sxz, pi, u, uipjp, ujpkp, uipkp);
            }
}

/* CPML corrections of the 9 velocity derivatives of a slab cell
 * (update_s_ssg_CPML_elastic.c:63-115, update_s_ssg_CPML.c:80-125) */
static void s_cpml(const og_grid *g, int j, int i, int k, vdiff *d, oreal *const prof[18], oreal *psi[9])
{
    const int fw = g->fw;
    const int ytop = !g->free_surf;
    oreal *p_vxx = psi[0], *p_vyx = psi[1], *p_vzx = psi[2];
    oreal *p_vxy = psi[3], *p_vyy = psi[4], *p_vzy = psi[5];
    oreal *p_vxz = psi[6], *p_vyz = psi[7], *p_vzz = psi[8];
    const int hx = slab_h(i, g->nx, fw, 1);
    const int hy = slab_h(j, g->ny, fw, ytop);
    const int hz = slab_h(k, g->nz, fw, 1);
    if (hx) {
        CPML_APPLY(p_vxx, psi_x_at(g, j, hx, k), d->vxx, PK(0, hx), PA(0, hx), PB(0, hx));
        CPML_APPLY(p_vyx, psi_x_at(g, j, hx, k), d->vyx, PKH(0, hx), PAH(0, hx), PBH(0, hx));
        CPML_APPLY(p_vzx, psi_x_at(g, j, hx, k), d->vzx, PKH(0, hx), PAH(0, hx), PBH(0, hx));
    }
    if (hy) {
        CPML_APPLY(p_vxy, psi_y_at(g, hy, i, k), d->vxy, PKH(1, hy), PAH(1, hy), PBH(1, hy));
        CPML_APPLY(p_vyy, psi_y_at(g, hy, i, k), d->vyy, PK(1, hy), PA(1, hy), PB(1, hy));
        CPML_APPLY(p_vzy, psi_y_at(g, hy, i, k), d->vzy, PKH(1, hy), PAH(1, hy), PBH(1, hy));
    }
    if (hz) {
        /* the low-z layer divides by the y-axis K profile
         * (update_s_ssg_CPML_elastic.c:104-108, update_s_ssg_CPML.c:110-116) */
        const int kax = (k <= fw) ? 1 : 2;
        CPML_APPLY(p_vxz, psi_z_at(g, j, i, hz), d->vxz, PKH(kax, hz), PAH(2, hz), PBH(2, hz));
        CPML_APPLY(p_vyz, psi_z_at(g, j, i, hz), d->vyz, PKH(kax, hz), PAH(2, hz), PBH(2, hz));
        CPML_APPLY(p_vzz, psi_z_at(g, j, i, hz), d->vzz, PK(kax, hz), PA(2, hz), PB(2, hz));
    }
}

void og_update_s_cpml_elastic(const og_grid *g, const int box[6],
                              const oreal *vx, const oreal *vy,
                              const oreal *vz, oreal *sxx, oreal *syy,
                              oreal *szz, oreal *sxy, oreal *syz, oreal *sxz,
                              const oreal *pi, const oreal *u,
                              const oreal *uipjp, const oreal *ujpkp,
                              const oreal *uipkp, oreal *const prof[18],
                              oreal *psi[9])
{
    oreal b1, b2;
    int i, j, k;
    vdiff d;
    fd4_coeffs(g->fdcoeff, &b1, &b2);
    #pragma omp parallel for private(i, k, d) schedule(static)
    for (j = 1; j <= g->ny; j++)
        for (i = 1; i <= g->nx; i++)
            for (k = 1; k <= g->nz; k++) {
                size_t c;
                if (in_box(box, j, i, k)) continue;
                c = og_idx(g, j, i, k);
                velocity_diffs(g, c, vx, vy, vz, b1, b2, &d);
                s_cpml(g, j, i, k, &d, prof, psi);
                stress_apply(g, c, &d, sxx, syy, szz, sxy, syz, sxz, pi, u, uipjp, ujpkp, uipkp);
            }
}

/* update_s_ssg.c:174-238: the relaxed stress step, the memory-variable
 * recursion, and the second half of the memory-variable term.  Written with
 * the reference's own float/double mix (1.0 + L*tau, 2.0*u, eta*0.5 are
 * double expressions rounded to float on assignment).  With L > 1 mechanisms
 * the memory variables of each component are summed in mechanism order
 * where the reference uses its one (the SLS sum of the tau-method), and
 * each mechanism is updated with its own eta; L = 1 is the reference's
 * arithmetic exactly. */
static inline void stress_apply_visco(const og_grid *g, size_t c, const vdiff *d,
                                      oreal *sxx, oreal *syy, oreal *szz,
                                      oreal *sxy, oreal *syz, oreal *sxz,
                                      const oreal *pi, const oreal *u,
                                      const oreal *uipjp, const oreal *ujpkp,
                                      const oreal *uipkp, const og_visco *V)
{
    const size_t fs = og_field_size(g);
    const oreal DT = g->dt;
    const oreal dthalbe = DT / 2.0;
    const int L = V->L;
    oreal fipjp, fjpkp, fipkp, gg, f, b, cc, dipjp, djpkp, dipkp, dd, e;
    oreal sr[6], *rc[6];
    int l, n;
    oreal vxyyx = d->vxy + d->vyx, vyzzy = d->vyz + d->vzy, vxzzx = d->vxz + d->vzx;
    oreal vxxyyzz = d->vxx + d->vyy + d->vzz;
    oreal vyyzz = d->vyy + d->vzz, vxxzz = d->vxx + d->vzz, vxxyy = d->vxx + d->vyy;
    for (n = 0; n < 6; n++) {
        rc[n] = V->r[n] + c;
        sr[n] = rc[n][0];
        for (l = 1; l < L; l++) sr[n] += rc[n][(size_t)l * fs];
    }
    fipjp = uipjp[c] * DT * (1.0 + L * V->tausipjp[c]);
    fjpkp = ujpkp[c] * DT * (1.0 + L * V->tausjpkp[c]);
    fipkp = uipkp[c] * DT * (1.0 + L * V->tausipkp[c]);
    gg = pi[c] * (1.0 + L * V->taup[c]);
    f = 2.0 * u[c] * (1.0 + L * V->taus[c]);
    /* sr: rxx ryy rzz rxy ryz rxz */
    sxy[c] += (fipjp * vxyyx) + (dthalbe * sr[3]);
    syz[c] += (fjpkp * vyzzy) + (dthalbe * sr[4]);
    sxz[c] += (fipkp * vxzzx) + (dthalbe * sr[5]);
    sxx[c] += DT * ((gg * vxxyyzz) - (f * vyyzz)) + (dthalbe * sr[0]);
    syy[c] += DT * ((gg * vxxyyzz) - (f * vxxzz)) + (dthalbe * sr[1]);
    szz[c] += DT * ((gg * vxxyyzz) - (f * vxxyy)) + (dthalbe * sr[2]);
    for (l = 0; l < L; l++) {
        const oreal eta = V->eta[l];
        const size_t o = (size_t)l * fs;
        b = 1.0 / (1.0 + (eta * 0.5));
        cc = 1.0 - (eta * 0.5);
        dipjp = uipjp[c] * eta * V->tausipjp[c];
        djpkp = ujpkp[c] * eta * V->tausjpkp[c];
        dipkp = uipkp[c] * eta * V->tausipkp[c];
        dd = 2.0 * u[c] * eta * V->taus[c];
        e = pi[c] * eta * V->taup[c];
        rc[3][o] = b * (rc[3][o] * cc - (dipjp * vxyyx));
        rc[4][o] = b * (rc[4][o] * cc - (djpkp * vyzzy));
        rc[5][o] = b * (rc[5][o] * cc - (dipkp * vxzzx));
        rc[0][o] = b * (rc[0][o] * cc - (e * vxxyyzz) + (dd * vyyzz));
        rc[1][o] = b * (rc[1][o] * cc - (e * vxxyyzz) + (dd * vxxzz));
        rc[2][o] = b * (rc[2][o] * cc - (e * vxxyyzz) + (dd * vxxyy));
    }
    for (n = 0; n < 6; n++) {
        sr[n] = rc[n][0];
        for (l = 1; l < L; l++) sr[n] += rc[n][(size_t)l * fs];
    }
    sxy[c] += (dthalbe * sr[3]);
    syz[c] += (dthalbe * sr[4]);
    sxz[c] += (dthalbe * sr[5]);
    sxx[c] += (dthalbe * sr[0]);
    syy[c] += (dthalbe * sr[1]);
    szz[c] += (dthalbe * sr[2]);
}

void og_update_s_visco(const og_grid *g, const int box[6], const oreal *vx,
                       const oreal *vy, const oreal *vz, oreal *sxx,
                       oreal *syy, oreal *szz, oreal *sxy, oreal *syz,
                       oreal *sxz, const oreal *pi, const oreal *u,
                       const oreal *uipjp, const oreal *ujpkp,
                       const oreal *uipkp, const og_visco *V)
{
    oreal b1, b2;
    int i, j, k;
    vdiff d;
    fd4_coeffs(g->fdcoeff, &b1, &b2);
    #pragma omp parallel for private(i, k, d) schedule(static)
    for (j = box[2]; j <= box[3]; j++)
        for (i = box[0]; i <= box[1]; i++)
            for (k = box[4]; k <= box[5]; k++) {
                size_t c = og_idx(g, j, i, k);
                velocity_diffs(g, c, vx, vy, vz, b1, b2, &d);
                stress_apply_visco(g, c, &d, sxx, syy, szz, sxy, syz, sxz, pi, u, uipjp, ujpkp, uipkp, V);
            }
}

void og_update_s_cpml_visco(const og_grid *g, const int box[6],
                            const oreal *vx, const oreal *vy,
                            const oreal *vz, oreal *sxx, oreal *syy,
                            oreal *szz, oreal *sxy, oreal *syz, oreal *sxz,
                            const oreal *pi, const oreal *u,
                            const oreal *uipjp, const oreal *ujpkp,
                            const oreal *uipkp, oreal *const prof[18],
                            oreal *psi[9], const og_visco *V)
{
    oreal b1, b2;
    int i, j, k;
    vdiff d;
    fd4_coeffs(g->fdcoeff, &b1, &b2);
    #pragma omp parallel for private(i, k, d) schedule(static)
    for (j = 1; j <= g->ny; j++)
        for (i = 1; i <= g->nx; i++)
            for (k = 1; k <= g->nz; k++) {
                size_t c;
                if (in_box(box, j, i, k)) continue;
                c = og_idx(g, j, i, k);
                velocity_diffs(g, c, vx, vy, vz, b1, b2, &d);
                s_cpml(g, j, i, k, &d, prof, psi);
                stress_apply_visco(g, c, &d, sxx, syy, szz, sxy, syz, sxz, pi, u, uipjp, ujpkp, uipkp, V);
            }
}

/* ------------------------------------------------------------------ */
/* psource.c:26-54 */
void og_psource(const og_grid *g, int nt, oreal *sxx, oreal *syy, oreal *szz,
                const og_source *srcs, const float *signals, int nsrc)
{
    int l;
    float amp = 0.0f;
    for (l = 0; l < nsrc; l++) {
        const float *s = signals + (size_t)l * (size_t)(g->nt + 2);
        size_t c;
        if (srcs[l].type != 1) continue;
        c = og_idx(g, srcs[l].j, srcs[l].i, srcs[l].k);
        if (nt == 1) amp = (float)(s[nt + 1] / (2.0 * g->dx * g->dy * g->dz));
        if ((nt > 1) && (nt < g->nt)) amp = (float)((s[nt + 1] - s[nt - 1]) / (2.0 * g->dx * g->dy * g->dz));
        if (nt == g->nt) amp = (float)(-s[nt - 1] / (2.0 * g->dx * g->dy * g->dz));
        sxx[c] += amp;
        syy[c] += amp;
        szz[c] += amp;
    }
}

/* surface_ssg_elastic.c:26-112, 4th order, surface at j = 1 */
void og_surface_elastic(const og_grid *g, const oreal *u, const oreal *pi,
                        oreal *sxx, oreal *syy, oreal *szz, oreal *sxy,
                        oreal *syz, const oreal *vx, const oreal *vy,
                        const oreal *vz)
{
    const ptrdiff_t sj = SJ(g), si = SI(g);
    const oreal DT = g->dt, DX = g->dx, DY = g->dy, DZ = g->dz;
    oreal b1, b2;
    int i, k, m;
    fd4_coeffs(g->fdcoeff, &b1, &b2);
    for (k = 1; k <= g->nz; k++)
        for (i = 1; i <= g->nx; i++) {
            size_t c = og_idx(g, 1, i, k);
            oreal vxx, vyy, vzz, f, gg, h;
            syy[c] = 0.0;
            for (m = 1; m <= 2; m++) {
                syy[c - m * sj] = -syy[c + m * sj];
                sxy[c - m * sj] = -sxy[c + (m - 1) * sj];
                syz[c - m * sj] = -syz[c + (m - 1) * sj];
            }
            vxx = DBWD(vx, c, si) / DX;
            vyy = DBWD(vy, c, sj) / DY;
            vzz = DBWD(vz, c, 1) / DZ;
            f = (oreal)(u[c] * 2.0);
            gg = pi[c];
            h = -(DT * (gg - f) * (gg - f) * (vxx + vzz) / gg) - (DT * (gg - f) * vyy);
            sxx[c] += h;
            szz[c] += h;
        }
}

/* seismo_ssg.c:27-53, SEISMO = 1 */
void og_seismo(const og_grid *g, int ins, int ntr, const int *rec_ijk,
               float *sec_x, float *sec_y, float *sec_z, int ns,
               const oreal *vx, const oreal *vy, const oreal *vz)
{
    int r;
    for (r = 0; r < ntr; r++) {
        size_t c = og_idx(g, rec_ijk[3 * r + 1], rec_ijk[3 * r + 0], rec_ijk[3 * r + 2]);
        size_t s = (size_t)r * (size_t)(ns + 1) + (size_t)ins;
        sec_x[s] = (float)vx[c];
        sec_y[s] = (float)vy[c];
        sec_z[s] = (float)vz[c];
    }
}

/* residual.c:27-56 (rt = 0 branch; the L2f diagnostic is not on the path) */
void og_residual(const og_grid *g, const float *obs, const float *syn,
                 float *diff, int ntr, int ns, float *l2)
{
    int r, j;
    for (r = 0; r < ntr; r++) {
        const float *o = obs + (size_t)r * (size_t)(ns + 1);
        const float *s = syn + (size_t)r * (size_t)(ns + 1);
        float *d = diff + (size_t)r * (size_t)(ns + 1);
        float vdiff = 0.0f;
        for (j = 1; j <= ns; j++) {
            vdiff += s[j] - o[j];
            d[ns - j + 1] = g->dt * vdiff;
            *l2 = (float)(*l2 + fabs(d[ns - j + 1] * d[ns - j + 1]));
        }
    }
}

/* disc_fourier.c:28-62 */
void og_discfourier(const og_grid *g, int nt, const oreal *vx,
                    const oreal *vy, const oreal *vz, oreal *fvx, oreal *fvy,
                    oreal *fvz, oreal *fivx, oreal *fivy, oreal *fivz,
                    const float *finv, int nf, int ntast, int back)
{
    const size_t fs = og_field_size(g);
    float t = 0.0f;
    int l, i, j, k;
    if (back == 0) t = nt * g->dt;
    if (back == 1) t = g->time - nt * g->dt;
    for (l = 1; l <= nf; l++) {
        double trig1 = cos(2.0 * t * finv[l - 1] * M_PI) * g->dt * ntast;
        double trig2 = sin(2.0 * t * finv[l - 1] * M_PI) * g->dt * ntast;
        size_t off = (size_t)(l - 1) * fs;
        #pragma omp parallel for private(i, k) schedule(static)
        for (j = 1; j <= g->ny; j++)
            for (i = 1; i <= g->nx; i++)
                for (k = 1; k <= g->nz; k++) {
                    size_t c = og_idx(g, j, i, k);
                    fvx[off + c] = (oreal)(fvx[off + c] + vx[c] * trig1);
                    fvy[off + c] = (oreal)(fvy[off + c] + vy[c] * trig1);
                    fvz[off + c] = (oreal)(fvz[off + c] + vz[c] * trig1);
                    fivx[off + c] = (oreal)(fivx[off + c] + vx[c] * trig2);
                    fivy[off + c] = (oreal)(fivy[off + c] + vy[c] * trig2);
                    fivz[off + c] = (oreal)(fivz[off + c] + vz[c] * trig2);
                }
    }
}

/* gradient_F.c:30-170 */
typedef struct { oreal xx, xy, xz, yx, yy, yz, zx, zy, zz; } grad9;

static inline void fgrad9(const og_grid *g, size_t c, const oreal *fx,
                          const oreal *fy, const oreal *fz, oreal b1, oreal b2,
                          grad9 *o)
{
    vdiff d;
    velocity_diffs(g, c, fx, fy, fz, b1, b2, &d);
    o->xx = d.vxx; o->xy = d.vxy; o->xz = d.vxz;
    o->yx = d.vyx; o->yy = d.vyy; o->yz = d.vyz;
    o->zx = d.vzx; o->zy = d.vzy; o->zz = d.vzz;
}

void og_gradient_F(const og_grid *g, const oreal *fvx, const oreal *fvy,
                   const oreal *fvz, const oreal *fivx, const oreal *fivy,
                   const oreal *fivz, const oreal *bvx, const oreal *bvy,
                   const oreal *bvz, const oreal *bivx, const oreal *bivy,
                   const oreal *bivz, oreal *grad1, oreal *grad2,
                   oreal *grad3, const oreal *rho, const oreal *pi,
                   const oreal *u, const float *finv, int nf)
{
    const size_t fs = og_field_size(g);
    oreal b1, b2;
    int l, i, j, k;
    fd4_coeffs(g->fdcoeff, &b1, &b2);
    for (l = 1; l <= nf; l++) {
        size_t off = (size_t)(l - 1) * fs;
        oreal fdummy = (oreal)(2.0 * 2.0 * finv[l - 1] * finv[l - 1] * M_PI * M_PI);
        fdummy = 1 / fdummy;
        #pragma omp parallel for private(i, k) schedule(static)
        for (j = 1; j <= g->ny; j++)
            for (i = 1; i <= g->nx; i++)
                for (k = 1; k <= g->nz; k++) {
                    size_t c = og_idx(g, j, i, k), co = off + c;
                    grad9 f, b, fi, bi;
                    oreal gradlam, gradmu, gradrho;
                    fgrad9(g, co, fvx, fvy, fvz, b1, b2, &f);
                    fgrad9(g, co, bvx, bvy, bvz, b1, b2, &b);
                    fgrad9(g, co, fivx, fivy, fivz, b1, b2, &fi);
                    fgrad9(g, co, bivx, bivy, bivz, b1, b2, &bi);
                    gradlam = (f.xx + f.yy + f.zz) * (b.xx + b.yy + b.zz) +
                              (fi.xx + fi.yy + fi.zz) * (bi.xx + bi.yy + bi.zz);
                    gradlam = -gradlam * fdummy;
                    gradmu = 2 * f.xx * b.xx + 2 * f.yy * b.yy + 2 * f.zz * b.zz +
                             2 * fi.xx * bi.xx + 2 * fi.yy * bi.yy + 2 * fi.zz * bi.zz +
                             (f.xy + f.yx) * (b.xy + b.yx) + (fi.xy + fi.yx) * (bi.xy + bi.yx) +
                             (f.xz + f.zx) * (b.xz + b.zx) + (fi.xz + fi.zx) * (bi.xz + bi.zx) +
                             (f.yz + f.zy) * (b.yz + b.zy) + (fi.yz + fi.zy) * (bi.yz + bi.zy);
                    gradmu = -gradmu * fdummy;
                    gradrho = (fvx[co] * bvx[co] + fvy[co] * bvy[co] + fvz[co] * bvz[co]) +
                              (fivx[co] * bivx[co] + fivy[co] * bivy[co] + fivz[co] * bivz[co]);
                    grad1[c] = (oreal)(grad1[c] + sqrt(rho[c] * pi[c]) * 2 * gradlam);
                    grad2[c] = (oreal)(grad2[c] + (-4 * sqrt(rho[c] * u[c]) * gradlam +
                                                   2 * sqrt(rho[c] * u[c]) * gradmu));
                    grad3[c] += gradrho + u[c] / rho[c] * gradmu + (pi[c] - 2 * u[c]) / rho[c] * gradlam;
                }
    }
}

/* hess_F.c:28-192: the diagonal Hessian (approximate, pseudo-Hessian of the
 * receiver-side Green's functions).  Forward fields f* of frequency l at
 * (l-1)*fs; receiver-side fields b* of receiver n (1..ntr_hess) and
 * frequency l at (n*nf + l - 1)*fs (hess_F.c:74); accumulated over n, then
 * l, cell by cell, in the reference's float expression order. */
void og_hess_F(const og_grid *g, const oreal *fvx, const oreal *fvy,
               const oreal *fvz, const oreal *fivx, const oreal *fivy,
               const oreal *fivz, const oreal *bvx, const oreal *bvy,
               const oreal *bvz, const oreal *bivx, const oreal *bivy,
               const oreal *bivz, oreal *hess1, oreal *hess2, oreal *hess3,
               const oreal *rho, const oreal *pi, const oreal *u,
               const float *finv, int nf, int ntr_hess)
{
    const size_t fs = og_field_size(g);
    oreal b1, b2;
    int n, l, i, j, k;
    fd4_coeffs(g->fdcoeff, &b1, &b2);
    for (n = 1; n <= ntr_hess; n++)
        for (l = 1; l <= nf; l++) {
            const int m = n * nf + l;
            const size_t off = (size_t)(l - 1) * fs, offb = (size_t)(m - 1) * fs;
            oreal fdummy = (oreal)(finv[l - 1] * M_PI * 2); /* hess_F.c:75-77 */
            fdummy = 1 / fdummy;
            #pragma omp parallel for private(i, k) schedule(static)
            for (j = 1; j <= g->ny; j++)
                for (i = 1; i <= g->nx; i++)
                    for (k = 1; k <= g->nz; k++) {
                        const size_t c = og_idx(g, j, i, k), co = off + c, cb = offb + c;
                        grad9 f, b, fi, bi;
                        oreal relam, imlam, remu, immu, rerho, imrho, revs, imvs, rerho1, imrho1;
                        fgrad9(g, co, fvx, fvy, fvz, b1, b2, &f);
                        fgrad9(g, cb, bvx, bvy, bvz, b1, b2, &b);
                        fgrad9(g, co, fivx, fivy, fivz, b1, b2, &fi);
                        fgrad9(g, cb, bivx, bivy, bivz, b1, b2, &bi);
                        /* hess_F.c:135-174 */
                        relam = (fi.xx + fi.yy + fi.zz) * (b.xx + b.yy + b.zz) +
                                (f.xx + f.yy + f.zz) * (bi.xx + bi.yy + bi.zz);
                        relam = fdummy * relam;
                        imlam = (fi.xx + fi.yy + fi.zz) * (bi.xx + bi.yy + bi.zz) -
                                (f.xx + f.yy + f.zz) * (b.xx + b.yy + b.zz);
                        imlam = fdummy * imlam;
                        remu = 2 * fi.xx * b.xx + 2 * fi.yy * b.yy + 2 * fi.zz * b.zz +
                               2 * f.xx * bi.xx + 2 * f.yy * bi.yy + 2 * f.zz * bi.zz +
                               (fi.xy + fi.yx) * (b.xy + b.yx) + (f.xy + f.yx) * (bi.xy + bi.yx) +
                               (fi.xz + fi.zx) * (b.xz + b.zx) + (f.xz + f.zx) * (bi.xz + bi.zx) +
                               (fi.yz + fi.zy) * (b.yz + b.zy) + (f.yz + f.zy) * (bi.yz + bi.zy);
                        remu = fdummy * remu;
                        immu = 2 * fi.xx * bi.xx + 2 * fi.yy * bi.yy + 2 * fi.zz * bi.zz -
                               2 * f.xx * b.xx - 2 * f.yy * b.yy - 2 * f.zz * b.zz +
                               (fi.xy + fi.yx) * (bi.xy + bi.yx) - (f.xy + f.yx) * (b.xy + b.yx) +
                               (fi.xz + fi.zx) * (bi.xz + bi.zx) - (f.xz + f.zx) * (b.xz + b.zx) +
                               (fi.yz + fi.zy) * (bi.yz + bi.zy) - (f.yz + f.zy) * (b.yz + b.zy);
                        immu = fdummy * immu;
                        rerho = (fvx[co] * bvx[cb] + fvy[co] * bvy[cb] + fvz[co] * bvz[cb]) -
                                (fivx[co] * bivx[cb] + fivy[co] * bivy[cb] + fivz[co] * bivz[cb]);
                        imrho = (fvx[co] * bivx[cb] + fvy[co] * bivy[cb] + fvz[co] * bivz[cb]) +
                                (fivx[co] * bvx[cb] + fivy[co] * bvy[cb] + fivz[co] * bvz[cb]);
                        revs = -2 * relam + remu;
                        imvs = -2 * imlam + immu;
                        rerho1 = rerho + relam * (pi[c] - 2 * u[c]) / rho[c] + remu * u[c] / rho[c];
                        imrho1 = imrho + imlam * (pi[c] - 2 * u[c]) / rho[c] + immu * u[c] / rho[c];
                        /* hess_F.c:180-182 */
                        hess1[c] += (relam * relam + imlam * imlam) * rho[c] * pi[c] * 4;
                        hess2[c] += (revs * revs + imvs * imvs) * 4 * rho[c] * u[c];
                        hess3[c] += rerho1 * rerho1 + imrho1 * imrho1;
                    }
        }
}

/* ifos3d.c:745-750 */
int og_ntast(float fmax, int tast, float dt)
{
    float dummy = (1 / (fmax * tast * dt));
    int ntast = (int)dummy;
    if (!ntast) ntast = 1;
    return ntast;
}

/* ------------------------------------------------------------------ */
/* one gradient shot, ifos3d.c:712-975 */
void og_td_gradient(const og_grid *g, const oreal *vfwd, const oreal *vadj, const oreal *rho,
                    const oreal *pi, const oreal *u, oreal *grad1, oreal *grad2, oreal *grad3)
{
    const size_t fs = og_field_size(g);
    const int N = g->nt;
    const oreal dt = (oreal)g->dt;
    oreal b1, b2;
    oreal *uf[3], *ub[3];
    double *acc[3];
    size_t c;
    int n, m, q, i, j, k;
    fd4_coeffs(g->fdcoeff, &b1, &b2);
    for (q = 0; q < 3; q++) {
        uf[q] = (oreal *)calloc(fs, sizeof(oreal));
        ub[q] = (oreal *)calloc(fs, sizeof(oreal));
        acc[q] = (double *)calloc(fs, sizeof(double));
    }
    /* forward samples N and N - 1 have no adjoint partner but enter uf */
    for (n = N; n >= N - 1 && n >= 1; n--)
        for (q = 0; q < 3; q++)
            for (c = 0; c < fs; c++) uf[q][c] = uf[q][c] + vfwd[((size_t)(n - 1) * 3 + q) * fs + c] * dt;
    for (m = 1; m <= N - 2; m++) {
        const oreal *vf[3], *vb[3];
        n = N - 1 - m;
        for (q = 0; q < 3; q++) {
            vf[q] = vfwd + ((size_t)(n - 1) * 3 + q) * fs;
            vb[q] = vadj + ((size_t)(m - 1) * 3 + q) * fs;
            for (c = 0; c < fs; c++) {
                uf[q][c] = uf[q][c] + vf[q][c] * dt;
                ub[q][c] = ub[q][c] + vb[q][c] * dt;
            }
        }
        #pragma omp parallel for private(i, k, c) schedule(static)
        for (j = 1; j <= g->ny; j++)
            for (i = 1; i <= g->nx; i++)
                for (k = 1; k <= g->nz; k++) {
                    grad9 a, b;
                    oreal gl, gm, gr;
                    c = og_idx(g, j, i, k);
                    fgrad9(g, c, uf[0], uf[1], uf[2], b1, b2, &a);
                    fgrad9(g, c, ub[0], ub[1], ub[2], b1, b2, &b);
                    gl = (a.xx + a.yy + a.zz) * (b.xx + b.yy + b.zz);
                    gl = -gl;
                    gm = 2 * a.xx * b.xx + 2 * a.yy * b.yy + 2 * a.zz * b.zz + (a.xy + a.yx) * (b.xy + b.yx) +
                         (a.xz + a.zx) * (b.xz + b.zx) + (a.yz + a.zy) * (b.yz + b.zy);
                    gm = -gm;
                    gr = vf[0][c] * vb[0][c] + vf[1][c] * vb[1][c] + vf[2][c] * vb[2][c];
                    acc[0][c] += (double)gl * (double)dt;
                    acc[1][c] += (double)gm * (double)dt;
                    acc[2][c] += (double)gr * (double)dt;
                }
    }
    #pragma omp parallel for private(i, k, c) schedule(static)
    for (j = 1; j <= g->ny; j++)
        for (i = 1; i <= g->nx; i++)
            for (k = 1; k <= g->nz; k++) {
                const double r = rho[og_idx(g, j, i, k)], p = pi[og_idx(g, j, i, k)], mu = u[og_idx(g, j, i, k)];
                const double srp = sqrt(r * p), sru = sqrt(r * mu);
                double l, gm_;
                c = og_idx(g, j, i, k);
                l = acc[0][c];
                gm_ = acc[1][c];
                grad1[c] = (oreal)((double)grad1[c] + 2.0 * srp * l);
                grad2[c] = (oreal)((double)grad2[c] + (-4.0 * sru * l + 2.0 * sru * gm_));
                grad3[c] = (oreal)((double)grad3[c] + (acc[2][c] + mu / r * gm_ + (p - 2.0 * mu) / r * l));
            }
    for (q = 0; q < 3; q++) {
        free(uf[q]);
        free(ub[q]);
        free(acc[q]);
    }
}

int og_run_shot(const og_grid *g, og_shot_io *io)
{
    const size_t fs = og_field_size(g);
    const int nt_all = g->nt, ntr = io->ntr, nf = io->nf;
    const int fw = g->fw;
    const size_t npx = (size_t)g->ny * 2 * fw * g->nz, npy = (size_t)2 * fw * g->nx * g->nz,
                 npz = (size_t)g->ny * g->nx * 2 * fw;
    int box[6], n, nt, rc = 0;
    oreal *rho = NULL, *pi = NULL, *u = NULL;
    oreal *f[9], *avg[6], *psi_s[9], *psi_v[9], *prof[18], *alpha[6];
    oreal *F[12];
    float *sig = NULL, *dx = NULL, *dy = NULL, *dz = NULL;
    og_source *recsrc = NULL;
    int ntast;
    const size_t sl = (size_t)(nt_all + 2);

    const int visco = io->L > 0 && io->L <= OG_MAX_L && io->taus && io->taup;
    oreal *taus = NULL, *taup = NULL, *tavg[3] = {NULL, NULL, NULL}, *rmem[6] = {NULL};
    og_visco V;
    memset(&V, 0, sizeof V);
    og_interior_box(g, box);
    ntast = og_ntast(io->finv[nf - 1], io->tast, g->dt);
    if (visco) {
        taus = (oreal *)calloc(fs, sizeof(oreal));
        taup = (oreal *)calloc(fs, sizeof(oreal));
        for (n = 0; n < 3; n++) tavg[n] = (oreal *)calloc(fs, sizeof(oreal));
        for (n = 0; n < 6; n++) rmem[n] = (oreal *)calloc(fs * (size_t)io->L, sizeof(oreal));
    }

    rho = (oreal *)calloc(fs, sizeof(oreal));
    pi = (oreal *)calloc(fs, sizeof(oreal));
    u = (oreal *)calloc(fs, sizeof(oreal));
    memcpy(rho, io->rho, fs * sizeof(oreal));
    memcpy(pi, io->pi, fs * sizeof(oreal));
    memcpy(u, io->u, fs * sizeof(oreal));
    og_matcopy_wrap(g, rho);
    og_matcopy_wrap(g, pi);
    og_matcopy_wrap(g, u);
    for (n = 0; n < 9; n++) f[n] = (oreal *)calloc(fs, sizeof(oreal));
    for (n = 0; n < 6; n++) avg[n] = (oreal *)calloc(fs, sizeof(oreal));
    for (n = 0; n < 9; n++) {
        size_t sz = (n < 3) ? npx : (n < 6) ? npy : npz;
        psi_s[n] = (oreal *)calloc(sz, sizeof(oreal));
    }
    /* psi_v order vxx vyx vzx | vxy vyy vzy | vxz vyz vzz -> x, y, z slabs */
    for (n = 0; n < 9; n++) {
        size_t sz = (n < 3) ? npx : (n < 6) ? npy : npz;
        psi_v[n] = (oreal *)calloc(sz, sizeof(oreal));
    }
    for (n = 0; n < 18; n++) prof[n] = (oreal *)calloc((size_t)(2 * fw + 1), sizeof(oreal));
    for (n = 0; n < 6; n++) alpha[n] = (oreal *)calloc((size_t)(2 * fw + 1), sizeof(oreal));
    for (n = 0; n < 12; n++) F[n] = (oreal *)calloc(fs * (size_t)nf, sizeof(oreal));
    sig = (float *)calloc(sl, sizeof(float));
    dx = (float *)calloc((size_t)ntr * (size_t)(nt_all + 2), sizeof(float));
    dy = (float *)calloc((size_t)ntr * (size_t)(nt_all + 2), sizeof(float));
    dz = (float *)calloc((size_t)ntr * (size_t)(nt_all + 2), sizeof(float));
    recsrc = (og_source *)calloc((size_t)(ntr > 0 ? ntr : 1), sizeof(og_source));

    og_av_mat(g, rho, u, NULL, avg[0], avg[1], avg[2], NULL, NULL, NULL, avg[3], avg[4], avg[5]);
    if (visco) {
        /* ifos3d.c:445-457, matcopy of taus/taup, av_mat.c:56-60 */
        oreal *tsa[3];
        memcpy(taus, io->taus, fs * sizeof(oreal));
        memcpy(taup, io->taup, fs * sizeof(oreal));
        og_matcopy_wrap(g, taus);
        og_matcopy_wrap(g, taup);
        for (n = 0; n < 3; n++) tsa[n] = tavg[n];
        og_av_mat(g, rho, u, taus, avg[0], avg[1], avg[2], tsa[0], tsa[1], tsa[2], avg[3], avg[4], avg[5]);
        V.taus = taus;
        V.taup = taup;
        V.tausipjp = tavg[0];
        V.tausjpkp = tavg[1];
        V.tausipkp = tavg[2];
        for (n = 0; n < OG_MAX_L; n++) V.eta[n] = io->eta[n];
        V.L = io->L;
        for (n = 0; n < 6; n++) V.r[n] = rmem[n];
    }
    og_cpml_coeff(g, io->vppml, io->fpml, prof, alpha);
    og_wavelet(g, &io->src, 1, io->source_shape, sig);

    /* forward, ifos3d.c:754-814 */
    for (nt = 1; nt <= nt_all; nt++) {
        if (isnan((double)f[1][og_idx(g, g->ny / 2, g->nx / 2, g->nz / 2)])) { rc = -1; goto done; }
        og_update_v(g, box, nt, f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8],
                    avg[3], avg[4], avg[5], &io->src, sig, sig, sig, 1, 0);
        og_update_v_cpml(g, box, nt, f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8],
                         avg[3], avg[4], avg[5], &io->src, sig, sig, sig, 1, 0, prof, psi_s);
        if (visco) {
            og_update_s_visco(g, box, f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8],
                              pi, u, avg[0], avg[1], avg[2], &V);
            og_update_s_cpml_visco(g, box, f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8],
                                   pi, u, avg[0], avg[1], avg[2], prof, psi_v, &V);
        } else {
            og_update_s_elastic(g, box, f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8],
                                pi, u, avg[0], avg[1], avg[2]);
            og_update_s_cpml_elastic(g, box, f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8],
                                     pi, u, avg[0], avg[1], avg[2], prof, psi_v);
        }
        og_psource(g, nt, f[3], f[4], f[5], &io->src, sig, 1);
        if (io->vfwd)
            for (n = 0; n < 3; n++) memcpy(io->vfwd + ((size_t)(nt - 1) * 3 + n) * fs, f[n], fs * sizeof(oreal));
        if (io->want_adjoint && nt % ntast == 0)
            og_discfourier(g, nt, f[0], f[1], f[2], F[0], F[1], F[2], F[3], F[4], F[5],
                           io->finv, nf, ntast, 0);
        if (g->free_surf)
            og_surface_elastic(g, u, pi, f[3], f[4], f[5], f[6], f[7], f[0], f[1], f[2]);
        og_seismo(g, nt, ntr, io->rec_ijk, io->syn_x, io->syn_y, io->syn_z, nt_all, f[0], f[1], f[2]);
    }
    if (!io->want_adjoint) goto done;

    /* residuals, ifos3d.c:837-866 */
    if (io->obs_x) {
        og_residual(g, io->obs_x, io->syn_x, dx + 0, ntr, nt_all, io->l2);
        og_residual(g, io->obs_y, io->syn_y, dy + 0, ntr, nt_all, io->l2);
        og_residual(g, io->obs_z, io->syn_z, dz + 0, ntr, nt_all, io->l2);
    }
    /* residual writes [ntr][ns+1]; re-pitch to [ntr][nt+2] for injection */
    {
        float *tx = (float *)calloc((size_t)ntr * sl, sizeof(float));
        float *ty = (float *)calloc((size_t)ntr * sl, sizeof(float));
        float *tz = (float *)calloc((size_t)ntr * sl, sizeof(float));
        int r;
        for (r = 0; r < ntr; r++) {
            memcpy(tx + (size_t)r * sl, dx + (size_t)r * (size_t)(nt_all + 1), sizeof(float) * (size_t)(nt_all + 1));
            memcpy(ty + (size_t)r * sl, dy + (size_t)r * (size_t)(nt_all + 1), sizeof(float) * (size_t)(nt_all + 1));
            memcpy(tz + (size_t)r * sl, dz + (size_t)r * (size_t)(nt_all + 1), sizeof(float) * (size_t)(nt_all + 1));
            recsrc[r].i = io->rec_ijk[3 * r + 0];
            recsrc[r].j = io->rec_ijk[3 * r + 1];
            recsrc[r].k = io->rec_ijk[3 * r + 2];
        }
        free(dx); free(dy); free(dz);
        dx = tx; dy = ty; dz = tz;
    }

    /* adjoint, ifos3d.c:905-960: fresh wavefield, receivers inject */
    for (n = 0; n < 9; n++) memset(f[n], 0, fs * sizeof(oreal));
    if (visco)
        for (n = 0; n < 6; n++) memset(rmem[n], 0, fs * (size_t)io->L * sizeof(oreal));
    for (n = 0; n < 9; n++) {
        size_t sz = (n < 3) ? npx : (n < 6) ? npy : npz;
        memset(psi_s[n], 0, sz * sizeof(oreal));
        memset(psi_v[n], 0, sz * sizeof(oreal));
    }
    for (nt = 1; nt <= nt_all; nt++) {
        if (isnan((double)f[1][og_idx(g, g->ny / 2, g->nx / 2, g->nz / 2)])) { rc = -1; goto done; }
        og_update_v(g, box, nt, f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8],
                    avg[3], avg[4], avg[5], recsrc, dx, dy, dz, ntr, 1);
        og_update_v_cpml(g, box, nt, f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8],
                         avg[3], avg[4], avg[5], recsrc, dx, dy, dz, ntr, 1, prof, psi_s);
        if (visco) {
            og_update_s_visco(g, box, f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8],
                              pi, u, avg[0], avg[1], avg[2], &V);
            og_update_s_cpml_visco(g, box, f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8],
                                   pi, u, avg[0], avg[1], avg[2], prof, psi_v, &V);
        } else {
            og_update_s_elastic(g, box, f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8],
                                pi, u, avg[0], avg[1], avg[2]);
            og_update_s_cpml_elastic(g, box, f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8],
                                     pi, u, avg[0], avg[1], avg[2], prof, psi_v);
        }
        if (io->vadj)
            for (n = 0; n < 3; n++) memcpy(io->vadj + ((size_t)(nt - 1) * 3 + n) * fs, f[n], fs * sizeof(oreal));
        if (nt % ntast == 0)
            og_discfourier(g, nt, f[0], f[1], f[2], F[6], F[7], F[8], F[9], F[10], F[11],
                           io->finv, nf, ntast, 1);
        if (g->free_surf)
            og_surface_elastic(g, u, pi, f[3], f[4], f[5], f[6], f[7], f[0], f[1], f[2]);
    }
    og_gradient_F(g, F[0], F[1], F[2], F[3], F[4], F[5], F[6], F[7], F[8], F[9], F[10], F[11],
                  io->grad1, io->grad2, io->grad3, rho, pi, u, io->finv, nf);
    if (io->ffwd)
        for (n = 0; n < 6; n++) memcpy(io->ffwd + (size_t)n * fs * nf, F[n], fs * nf * sizeof(oreal));
    if (io->fadj)
        for (n = 0; n < 6; n++) memcpy(io->fadj + (size_t)n * fs * nf, F[6 + n], fs * nf * sizeof(oreal));

done:
    free(rho); free(pi); free(u);
    for (n = 0; n < 9; n++) { free(f[n]); free(psi_s[n]); free(psi_v[n]); }
    for (n = 0; n < 6; n++) { free(avg[n]); free(alpha[n]); }
    for (n = 0; n < 18; n++) free(prof[n]);
    for (n = 0; n < 12; n++) free(F[n]);
    free(sig); free(dx); free(dy); free(dz); free(recsrc);
    free(taus); free(taup);
    for (n = 0; n < 3; n++) free(tavg[n]);
    for (n = 0; n < 6; n++) free(rmem[n]);
    return rc;
}

/* ------------------------------------------------------------------ */
/* The consistent adjoint (ifos_oracle.h): the transpose of og_run_shot's
 * forward step, run backwards over a stored forward history. */

/* derivative components of the velocity update, in sdiff order
 * (sxx_x sxy_y sxz_z syy_y sxy_x syz_z szz_z sxz_x syz_y): differentiated
 * field, axis, forward (1) or backward (0) difference, half-grid CPML
 * profile, CPML memory array (og_update_v_cpml) */
static const int VQ_FIELD[9] = {3, 6, 8, 4, 6, 7, 5, 8, 7};
static const int VQ_AXIS[9] = {0, 1, 2, 1, 0, 2, 2, 0, 1};
static const int VQ_FWD[9] = {1, 0, 0, 1, 0, 0, 1, 0, 0};
static const int VQ_HALF[9] = {1, 0, 0, 1, 0, 0, 1, 0, 0};
static const int VQ_PSI[9] = {0, 3, 6, 4, 1, 7, 8, 2, 5};
/* the same for the stress update, vdiff order (vxx vxy vxz vyx vyy vyz vzx
 * vzy vzz; s_cpml) */
static const int SQ_FIELD[9] = {0, 0, 0, 1, 1, 1, 2, 2, 2};
static const int SQ_AXIS[9] = {0, 1, 2, 0, 1, 2, 0, 1, 2};
static const int SQ_FWD[9] = {0, 1, 1, 1, 0, 1, 1, 1, 0};
static const int SQ_HALF[9] = {0, 1, 1, 1, 0, 1, 1, 1, 0};
static const int SQ_PSI[9] = {0, 3, 6, 1, 4, 7, 2, 5, 8};

/* CPML profile values and memory-variable slot of one derivative component
 * at (j, i, k); 0 where the component has no memory variable.  stress: the
 * z layers take K from the y axis at the low side (s_cpml). */
static int q_cpml(const og_grid *g, oreal *const prof[18], int stress, int q, int j, int i, int k, oreal *K,
                  oreal *a, oreal *b, int *arr, size_t *at)
{
    const int fw = g->fw;
    const int ax = stress ? SQ_AXIS[q] : VQ_AXIS[q];
    const int half = stress ? SQ_HALF[q] : VQ_HALF[q];
    const int h = ax == 0 ? slab_h(i, g->nx, fw, 1) : ax == 1 ? slab_h(j, g->ny, fw, !g->free_surf)
                                                                : slab_h(k, g->nz, fw, 1);
    int kax = ax;
    if (!h) return 0;
    if (stress && ax == 2) kax = (k <= fw) ? 1 : 2;
    *K = half ? PKH(kax, h) : PK(kax, h);
    *a = half ? PAH(ax, h) : PA(ax, h);
    *b = half ? PBH(ax, h) : PB(ax, h);
    *arr = stress ? SQ_PSI[q] : VQ_PSI[q];
    *at = ax == 0 ? psi_x_at(g, j, h, k) : ax == 1 ? psi_y_at(g, h, i, k) : psi_z_at(g, j, i, h);
    return 1;
}

/* transposed 4th-order differences at x (L zero outside [1..N]^3):
 * DFWD^T L = b1 (L[x-s] - L[x]) + b2 (L[x-2s] - L[x+s]),
 * DBWD^T L = b1 (L[x] - L[x+s]) + b2 (L[x-s] - L[x+2s]) */
static inline oreal dT(const oreal *L, size_t x, ptrdiff_t s, int fwd, oreal b1, oreal b2)
{
    if (fwd) return b1 * (L[x - s] - L[x]) + b2 * (L[x - 2 * s] - L[x + s]);
    return b1 * (L[x] - L[x + s]) + b2 * (L[x - s] - L[x + 2 * s]);
}

static inline ptrdiff_t axis_stride(const og_grid *g, int ax)
{
    return ax == 0 ? SI(g) : ax == 1 ? SJ(g) : 1;
}

int og_run_shot_adj(const og_grid *g, og_adj_io *io)
{
    const size_t fs = og_field_size(g);
    const int N = g->nt, ntr = io->ntr, fw = g->fw;
    const size_t npx = (size_t)g->ny * 2 * fw * g->nz, npy = (size_t)2 * fw * g->nx * g->nz,
                 npz = (size_t)g->ny * g->nx * 2 * fw;
    const size_t np[9] = {npx, npx, npx, npy, npy, npy, npz, npz, npz};
    const ptrdiff_t sj = SJ(g), si = SI(g);
    const oreal DT = (oreal)g->dt;
    const oreal dtd[3] = {(oreal)g->dt / (oreal)g->dx, (oreal)g->dt / (oreal)g->dy, (oreal)g->dt / (oreal)g->dz};
    const oreal rh[3] = {(oreal)1 / (oreal)g->dx, (oreal)1 / (oreal)g->dy, (oreal)1 / (oreal)g->dz};
    oreal b1, b2;
    int box[6], n, q, nt, i, j, k, rc = 0;
    oreal *rho, *pi, *u, *avg[6], *prof[18], *alpha[6];
    oreal **hf, **hps, **hpv;   /* history: [(N+1)*9] fields, psi_s, psi_v */
    oreal *L[9], *mus[9], *muv[9], *scr[9];
    double *G[8];               /* rip rjp rkp pi u uipjp ujpkp uipkp */
    double *Gr, *Gp, *Gu;
    float *sig;
    float *syn[3] = {io->syn_x, io->syn_y, io->syn_z};
    const float *obs[3] = {io->obs_x, io->obs_y, io->obs_z};
    oreal *sr[3];  /* the seismograms in the oracle's own precision: J and the residuals come from them */
    double J = 0.0;

    if (g->free_surf || io->src.type != 1 || g->fdcoeff < 1) return -2;
    fd4_coeffs(g->fdcoeff, &b1, &b2);
    og_interior_box(g, box);
    rho = (oreal *)calloc(fs, sizeof(oreal));
    pi = (oreal *)calloc(fs, sizeof(oreal));
    u = (oreal *)calloc(fs, sizeof(oreal));
    memcpy(rho, io->rho, fs * sizeof(oreal));
    memcpy(pi, io->pi, fs * sizeof(oreal));
    memcpy(u, io->u, fs * sizeof(oreal));
    og_matcopy_wrap(g, rho);
    og_matcopy_wrap(g, pi);
    og_matcopy_wrap(g, u);
    for (n = 0; n < 6; n++) avg[n] = (oreal *)calloc(fs, sizeof(oreal));
    og_av_mat(g, rho, u, NULL, avg[0], avg[1], avg[2], NULL, NULL, NULL, avg[3], avg[4], avg[5]);
    for (n = 0; n < 18; n++) prof[n] = (oreal *)calloc((size_t)(2 * fw + 1), sizeof(oreal));
    for (n = 0; n < 6; n++) alpha[n] = (oreal *)calloc((size_t)(2 * fw + 1), sizeof(oreal));
    og_cpml_coeff(g, io->vppml, io->fpml, prof, alpha);
    sig = (float *)calloc((size_t)(N + 2), sizeof(float));
    og_wavelet(g, &io->src, 1, io->source_shape, sig);

    for (q = 0; q < 3; q++) sr[q] = (oreal *)calloc((size_t)(ntr > 0 ? ntr : 1) * (size_t)(N + 1), sizeof(oreal));
    hf = (oreal **)calloc((size_t)(N + 1) * 9, sizeof(oreal *));
    hps = (oreal **)calloc((size_t)(N + 1) * 9, sizeof(oreal *));
    hpv = (oreal **)calloc((size_t)(N + 1) * 9, sizeof(oreal *));
    for (n = 0; n < 9; n++) {
        hf[n] = (oreal *)calloc(fs, sizeof(oreal));
        hps[n] = (oreal *)calloc(np[n], sizeof(oreal));
        hpv[n] = (oreal *)calloc(np[n], sizeof(oreal));
    }
    /* forward (og_run_shot's steps without the DFT), state kept after every step */
    for (nt = 1; nt <= N; nt++) {
        oreal **f = hf + (size_t)nt * 9, **ps = hps + (size_t)nt * 9, **pv = hpv + (size_t)nt * 9;
        for (n = 0; n < 9; n++) {
            f[n] = (oreal *)malloc(fs * sizeof(oreal));
            ps[n] = (oreal *)malloc(np[n] * sizeof(oreal));
            pv[n] = (oreal *)malloc(np[n] * sizeof(oreal));
            memcpy(f[n], hf[(size_t)(nt - 1) * 9 + n], fs * sizeof(oreal));
            memcpy(ps[n], hps[(size_t)(nt - 1) * 9 + n], np[n] * sizeof(oreal));
            memcpy(pv[n], hpv[(size_t)(nt - 1) * 9 + n], np[n] * sizeof(oreal));
        }
        if (isnan((double)f[1][og_idx(g, g->ny / 2, g->nx / 2, g->nz / 2)])) { rc = -1; break; }
        og_update_v(g, box, nt, f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8],
                    avg[3], avg[4], avg[5], &io->src, sig, sig, sig, 1, 0);
        og_update_v_cpml(g, box, nt, f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8],
                         avg[3], avg[4], avg[5], &io->src, sig, sig, sig, 1, 0, prof, ps);
        og_update_s_elastic(g, box, f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8],
                            pi, u, avg[0], avg[1], avg[2]);
        og_update_s_cpml_elastic(g, box, f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8],
                                 pi, u, avg[0], avg[1], avg[2], prof, pv);
        og_psource(g, nt, f[3], f[4], f[5], &io->src, sig, 1);
        og_seismo(g, nt, ntr, io->rec_ijk, io->syn_x, io->syn_y, io->syn_z, N, f[0], f[1], f[2]);
        for (n = 0; n < ntr; n++) {
            const int *p = io->rec_ijk + 3 * n;
            const size_t c = og_idx(g, p[1], p[0], p[2]), o = (size_t)n * (size_t)(N + 1) + (size_t)nt;
            for (q = 0; q < 3; q++) sr[q][o] = f[q][c];
        }
    }
    if (rc) goto done;
    /* misfit */
    for (q = 0; q < 3; q++)
        for (n = 0; n < ntr; n++)
            for (nt = 1; nt <= N; nt++) {
                const size_t o = (size_t)n * (size_t)(N + 1) + (size_t)nt;
                const double r = (double)sr[q][o] - (double)obs[q][o];
                J += 0.5 * r * r;
            }
    *io->J = J;

    for (n = 0; n < 9; n++) {
        L[n] = (oreal *)calloc(fs, sizeof(oreal));
        scr[n] = (oreal *)calloc(fs, sizeof(oreal));
        mus[n] = (oreal *)calloc(np[n], sizeof(oreal));
        muv[n] = (oreal *)calloc(np[n], sizeof(oreal));
    }
    for (n = 0; n < 8; n++) G[n] = (double *)calloc(fs, sizeof(double));

    for (nt = N; nt >= 1; nt--) {
        oreal *const *f = hf + (size_t)nt * 9, *const *fp = hf + (size_t)(nt - 1) * 9;
        oreal *const *ps = hps + (size_t)nt * 9, *const *pv = hpv + (size_t)nt * 9;
        /* receivers: d syn(nt) / d v(nt) */
        for (n = 0; n < ntr; n++) {
            const int *p = io->rec_ijk + 3 * n;
            const size_t c = og_idx(g, p[1], p[0], p[2]), o = (size_t)n * (size_t)(N + 1) + (size_t)nt;
            for (q = 0; q < 3; q++) L[q][c] += (oreal)((double)sr[q][o] - (double)obs[q][o]);
        }
        /* stress update of step nt: sigma(nt) = sigma(nt-1) + M(model) e'(v(nt)) */
#pragma omp parallel for private(i, k, q) schedule(static)
        for (j = 1; j <= g->ny; j++)
            for (i = 1; i <= g->nx; i++)
                for (k = 1; k <= g->nz; k++) {
                    const size_t c = og_idx(g, j, i, k);
                    const oreal Lxx = L[3][c], Lyy = L[4][c], Lzz = L[5][c], Lxy = L[6][c], Lyz = L[7][c],
                                Lxz = L[8][c];
                    const oreal gg = pi[c], f2 = (oreal)(2.0 * u[c]), sumL = Lxx + Lyy + Lzz;
                    oreal gq[9], e[9];
                    vdiff d;
                    gq[0] = DT * (gg * sumL - f2 * (Lyy + Lzz));
                    gq[4] = DT * (gg * sumL - f2 * (Lxx + Lzz));
                    gq[8] = DT * (gg * sumL - f2 * (Lxx + Lyy));
                    gq[1] = gq[3] = avg[0][c] * DT * Lxy;
                    gq[5] = gq[7] = avg[1][c] * DT * Lyz;
                    gq[2] = gq[6] = avg[2][c] * DT * Lxz;
                    /* the forward's CPML-corrected derivatives of v(nt) */
                    velocity_diffs(g, c, f[0], f[1], f[2], b1, b2, &d);
                    e[0] = d.vxx; e[1] = d.vxy; e[2] = d.vxz; e[3] = d.vyx; e[4] = d.vyy;
                    e[5] = d.vyz; e[6] = d.vzx; e[7] = d.vzy; e[8] = d.vzz;
                    for (q = 0; q < 9; q++) {
                        oreal K, a, b;
                        int arr;
                        size_t at;
                        if (q_cpml(g, prof, 1, q, j, i, k, &K, &a, &b, &arr, &at)) {
                            const oreal m = muv[arr][at] + gq[q];
                            e[q] = e[q] / K + pv[arr][at];
                            muv[arr][at] = b * m;
                            scr[q][c] = gq[q] / K + a * m;
                        } else {
                            scr[q][c] = gq[q];
                        }
                    }
                    G[3][c] += (double)DT * ((double)e[0] + e[4] + e[8]) * sumL;
                    G[4][c] += -2.0 * (double)DT * (((double)e[4] + e[8]) * Lxx + ((double)e[0] + e[8]) * Lyy +
                                                    ((double)e[0] + e[4]) * Lzz);
                    G[5][c] += (double)DT * ((double)e[1] + e[3]) * Lxy;
                    G[6][c] += (double)DT * ((double)e[5] + e[7]) * Lyz;
                    G[7][c] += (double)DT * ((double)e[2] + e[6]) * Lxz;
                }
#pragma omp parallel for private(i, k, q) schedule(static)
        for (j = 1; j <= g->ny; j++)
            for (i = 1; i <= g->nx; i++)
                for (k = 1; k <= g->nz; k++) {
                    const size_t c = og_idx(g, j, i, k);
                    for (q = 0; q < 9; q++)
                        L[SQ_FIELD[q]][c] += rh[SQ_AXIS[q]] * dT(scr[q], c, axis_stride(g, SQ_AXIS[q]), SQ_FWD[q], b1, b2);
                }
        /* velocity update of step nt: v(nt) = v(nt-1) + d'(sigma(nt-1)) / rho-bar */
#pragma omp parallel for private(i, k, q) schedule(static)
        for (j = 1; j <= g->ny; j++)
            for (i = 1; i <= g->nx; i++)
                for (k = 1; k <= g->nz; k++) {
                    const size_t c = og_idx(g, j, i, k);
                    const oreal hv[3] = {L[0][c] / avg[5][c], L[1][c] / avg[3][c], L[2][c] / avg[4][c]};
                    oreal dd[9];
                    sdiff d;
                    stress_diffs(g, c, fp[3], fp[4], fp[5], fp[6], fp[7], fp[8], b1, b2, &d);
                    dd[0] = d.sxx_x; dd[1] = d.sxy_y; dd[2] = d.sxz_z; dd[3] = d.syy_y; dd[4] = d.sxy_x;
                    dd[5] = d.syz_z; dd[6] = d.szz_z; dd[7] = d.sxz_x; dd[8] = d.syz_y;
                    for (q = 0; q < 9; q++) {
                        const oreal gqq = hv[q / 3];
                        oreal K, a, b;
                        int arr;
                        size_t at;
                        if (q_cpml(g, prof, 0, q, j, i, k, &K, &a, &b, &arr, &at)) {
                            const oreal m = mus[arr][at] + gqq;
                            dd[q] = dd[q] / K + ps[arr][at];
                            mus[arr][at] = b * m;
                            scr[q][c] = gqq / K + a * m;
                        } else {
                            scr[q][c] = gqq;
                        }
                    }
                    G[0][c] -= (double)L[0][c] * ((double)dd[0] + dd[1] + dd[2]) / ((double)avg[5][c] * avg[5][c]);
                    G[1][c] -= (double)L[1][c] * ((double)dd[3] + dd[4] + dd[5]) / ((double)avg[3][c] * avg[3][c]);
                    G[2][c] -= (double)L[2][c] * ((double)dd[6] + dd[7] + dd[8]) / ((double)avg[4][c] * avg[4][c]);
                }
#pragma omp parallel for private(i, k, q) schedule(static)
        for (j = 1; j <= g->ny; j++)
            for (i = 1; i <= g->nx; i++)
                for (k = 1; k <= g->nz; k++) {
                    const size_t c = og_idx(g, j, i, k);
                    for (q = 0; q < 9; q++)
                        L[VQ_FIELD[q]][c] += dtd[VQ_AXIS[q]] * dT(scr[q], c, axis_stride(g, VQ_AXIS[q]), VQ_FWD[q], b1, b2);
                }
    }

    /* av_mat's averages -> cells (rip = (rho + rho(i+1)) / 2, ..., uipjp =
     * 4 / sum 1/u over 4 cells), matcopy's halo copies -> their source
     * cells, then (pi, u, rho) -> (vp, vs, rho) */
    Gr = (double *)calloc(fs, sizeof(double));
    Gp = (double *)calloc(fs, sizeof(double));
    Gu = (double *)calloc(fs, sizeof(double));
    for (j = 1; j <= g->ny; j++)
        for (i = 1; i <= g->nx; i++)
            for (k = 1; k <= g->nz; k++) {
                const size_t c = og_idx(g, j, i, k);
                const ptrdiff_t nb[3][4] = {{0, si, sj + si, sj}, {0, 1, sj + 1, sj}, {0, 1, si + 1, si}};
                int m;
                Gp[c] += G[3][c];
                Gu[c] += G[4][c];
                Gr[c] += 0.5 * (G[0][c] + G[1][c] + G[2][c]);
                Gr[c + si] += 0.5 * G[0][c];
                Gr[c + sj] += 0.5 * G[1][c];
                Gr[c + 1] += 0.5 * G[2][c];
                for (m = 0; m < 3; m++) {
                    const double w = (double)avg[m][c] * (double)avg[m][c] * 0.25;
                    int t;
                    for (t = 0; t < 4; t++) {
                        const double ux = u[c + nb[m][t]];
                        Gu[c + nb[m][t]] += G[5 + m][c] * w / (ux * ux);
                    }
                }
            }
    for (j = 0; j <= g->ny + 1; j++)
        for (i = 0; i <= g->nx + 1; i++)
            for (k = 0; k <= g->nz + 1; k++) {
                const int wj = j == 0 ? g->ny : (j == g->ny + 1 ? 1 : j);
                const int wi = i == 0 ? g->nx : (i == g->nx + 1 ? 1 : i);
                const int wk = k == 0 ? g->nz : (k == g->nz + 1 ? 1 : k);
                const size_t c = og_idx(g, j, i, k), t = og_idx(g, wj, wi, wk);
                if (c == t) continue;
                Gr[t] += Gr[c];
                Gp[t] += Gp[c];
                Gu[t] += Gu[c];
                Gr[c] = Gp[c] = Gu[c] = 0.0;
            }
    for (j = 1; j <= g->ny; j++)
        for (i = 1; i <= g->nx; i++)
            for (k = 1; k <= g->nz; k++) {
                const size_t c = og_idx(g, j, i, k);
                const double r = rho[c], p = pi[c], m = u[c];
                io->grad1[c] = (oreal)((double)io->grad1[c] + 2.0 * sqrt(r * p) * Gp[c]);
                io->grad2[c] = (oreal)((double)io->grad2[c] + 2.0 * sqrt(r * m) * Gu[c]);
                io->grad3[c] = (oreal)((double)io->grad3[c] + Gr[c] + Gp[c] * p / r + Gu[c] * m / r);
            }
    free(Gr);
    free(Gp);
    free(Gu);
    for (n = 0; n < 9; n++) {
        free(L[n]);
        free(scr[n]);
        free(mus[n]);
        free(muv[n]);
    }
    for (n = 0; n < 8; n++) free(G[n]);
done:
    for (q = 0; q < 3; q++) free(sr[q]);
    for (n = 0; n < (N + 1) * 9; n++) {
        free(hf[n]);
        free(hps[n]);
        free(hpv[n]);
    }
    free(hf);
    free(hps);
    free(hpv);
    free(rho);
    free(pi);
    free(u);
    for (n = 0; n < 6; n++) {
        free(avg[n]);
        free(alpha[n]);
    }
    for (n = 0; n < 18; n++) free(prof[n]);
    free(sig);
    return rc;
}
