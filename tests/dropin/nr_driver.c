/*
 * nr_driver.c -- a caller of libifos3d_dropin.so shaped like the IFOS3D
 * driver (test harness).
 *
 * It owns the globals the routines read (names and types of IFOS3D's
 * globvar.h), allocates Numerical-Recipes tensors (one contiguous block,
 * pointer tables offset by the lower bounds, the layout of util.c:250-400),
 * and runs one gradient shot by calling the drop-in routines in the order of
 * ifos3d.c:712-975: zero_wavefield / zero_invers, forward steps (update_v,
 * update_v_CPML, update_s_elastic, update_s_CPML_elastic, psource,
 * discfourier, surface_elastic, seismo), residual (host, residual.c:39-53),
 * adjoint steps, gradient_F.  Model set-up follows ifos3d.c:681-683
 * (single-rank matcopy, then av_mat) and ifos3d.c:553-627 (CPML_ini_elastic,
 * CPML_coeff).  Inputs and outputs cross the ctypes boundary as flat arrays
 * in the canonical layout of the tests (halo 2, [j][i][k]).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "ifos3d_dropin.h"
#include "ifos3d_hip.h"

/* the globals the drop-in routines read */
float DT, DX, DY, DZ, TIME, ALPHA, BETA, VPPML, FPML;
int NT, NX, NY, NZ, FDORDER, FDCOEFF, ABS_TYPE, FW, FREE_SURF, L, SEISMO, NDT = 1, NDTSHIFT = 0;
int POS[4], NPROCX, NPROCY, NPROCZ;
int LBFGS, NUMPAR, BFGSNUM;
float VP0, VS0, RHO0, WEIGHT[3];

/* ---- Numerical-Recipes style tensors -------------------------------- */
static float ***nr3(int jl, int jh, int il, int ih, int kl, int kh)
{
    int nj = jh - jl + 1, ni = ih - il + 1, nk = kh - kl + 1, j, i;
    float *block = (float *)calloc((size_t)nj * ni * nk, sizeof(float));
    float **rows = (float **)malloc(sizeof(float *) * (size_t)nj * ni);
    float ***t = (float ***)malloc(sizeof(float **) * (size_t)nj);
    for (j = 0; j < nj; j++) {
        t[j] = rows + (size_t)j * ni - il;
        for (i = 0; i < ni; i++) rows[(size_t)j * ni + i] = block + ((size_t)j * ni + i) * nk - kl;
    }
    return t - jl;
}

static void nr3_free(float ***t, int jl, int il, int kl)
{
    free(&t[jl][il][kl]);
    free(&t[jl][il]);
    free(&t[jl]);
}

static float ****nr4(int ml, int mh, int jl, int jh, int il, int ih, int kl, int kh)
{
    int nm = mh - ml + 1, m;
    float ****t = (float ****)malloc(sizeof(float ***) * (size_t)nm);
    for (m = 0; m < nm; m++) t[m] = nr3(jl, jh, il, ih, kl, kh);
    return t - ml;
}

static void nr4_free(float ****t, int ml, int mh, int jl, int il, int kl)
{
    int m;
    for (m = ml; m <= mh; m++) nr3_free(t[m], jl, il, kl);
    free(&t[ml]);
}

static float **nr2(int rl, int rh, int cl, int ch)
{
    int nr = rh - rl + 1, r;
    float *block = (float *)calloc((size_t)nr * (ch - cl + 1), sizeof(float));
    float **m = (float **)malloc(sizeof(float *) * (size_t)nr);
    for (r = 0; r < nr; r++) m[r] = block + (size_t)r * (ch - cl + 1) - cl;
    return m - rl;
}

static void nr2_free(float **m, int rl, int cl)
{
    free(&m[rl][cl]);
    free(&m[rl]);
}

static size_t cidx(int j, int i, int k)
{
    return ((size_t)(j + 2) * (NX + 5) + (i + 2)) * (size_t)(NZ + 5) + (k + 2);
}

static void from_canon(float ***t, const float *c, int jl, int jh, int il, int ih, int kl, int kh)
{
    int j, i, k;
    for (j = jl; j <= jh; j++)
        for (i = il; i <= ih; i++)
            for (k = kl; k <= kh; k++) t[j][i][k] = c[cidx(j, i, k)];
}

static void to_canon(float ***t, float *c, int jl, int jh, int il, int ih, int kl, int kh)
{
    int j, i, k;
    for (j = jl; j <= jh; j++)
        for (i = il; i <= ih; i++)
            for (k = kl; k <= kh; k++) c[cidx(j, i, k)] = t[j][i][k];
}

/* single-rank matcopy: every halo face takes the opposite face (periodic) */
static void matcopy1(float ***m)
{
    int i, j, k;
    for (i = 0; i <= NX + 1; i++)
        for (k = 0; k <= NZ + 1; k++) {
            float top = m[1][i][k], bot = m[NY][i][k];
            m[NY + 1][i][k] = top;
            m[0][i][k] = bot;
        }
    for (j = 0; j <= NY + 1; j++)
        for (k = 0; k <= NZ + 1; k++) {
            float lef = m[j][1][k], rig = m[j][NX][k];
            m[j][NX + 1][k] = lef;
            m[j][0][k] = rig;
        }
    for (i = 0; i <= NX + 1; i++)
        for (j = 0; j <= NY + 1; j++) {
            float fro = m[j][i][1], bac = m[j][i][NZ];
            m[j][i][NZ + 1] = fro;
            m[j][i][0] = bac;
        }
}

/* residual.c:39-53 on the host (the drop-in leaves it to the caller) */
static void residual1(float **obs, float **syn, float **diff, int ntr, int ns, float *l2)
{
    int i, j;
    for (i = 1; i <= ntr; i++) {
        float vdiff = 0.0f;
        for (j = 1; j <= ns; j++) {
            vdiff += syn[i][j] - obs[i][j];
            diff[i][ns - j + 1] = DT * vdiff;
            *l2 = (float)(*l2 + fabs(diff[i][ns - j + 1] * diff[i][ns - j + 1]));
        }
    }
}

/*
 * One shot.  geometry: nx ny nz fw; dx dy dz dt; nt; fdcoeff; free_surf.
 * rho/pi/u: canonical arrays.  src: i j k type, tshift fc amp, shape.
 * rec_ijk [ntr][3].  obs [3][ntr][nt+1] (NULL: forward only).
 * finv[nf], tast, vppml, fpml.  Outputs: syn [3][ntr][nt+1], grads [3]
 * canonical (accumulated), *l2.  Returns 0, or -1 on the NaN probe.
 */
/* optional viscoelastic run: relaxation ratios (canonical) and eta; NULL
 * switches back to the elastic update */
/* seconds of the last dd_run_shot: [0] the shot's library calls
 * (zero_wavefield ... gradient_F, the per-step host copies included), [1]
 * the set-up before them (NR tensors, model, matcopy, av_mat, CPML), [2] the
 * whole call */
double dd_last_times[3];
static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static const float *g_taus, *g_taup;
static float g_eta;
void dd_set_visco(const float *taus_c, const float *taup_c, float eta)
{
    g_taus = taus_c;
    g_taup = taup_c;
    g_eta = eta;
}

int dd_run_shot(int nx, int ny, int nz, int fw, float dx, float dy, float dz, float dt, int nt, int fdcoeff,
                int free_surf, const float *rho_c, const float *pi_c, const float *u_c, const int *src,
                const float *src_f, int shape, int ntr, const int *rec_ijk, const float *obs, const float *finv,
                int nf, int tast, float vppml, float fpml, float *syn, float *grads, float *l2)
{
    int j, n, r, c, xb[2], yb[2], zb[2], ntast, rc = 0;
    const int lo = -2;
    float ***vx, ***vy, ***vz, ***sxx, ***syy, ***szz, ***sxy, ***syz, ***sxz;
    float ***rho, ***pi, ***u, ***uipjp, ***ujpkp, ***uipkp, ***rjp, ***rkp, ***rip;
    float ***g1, ***g2, ***g3, ***absorb;
    float ***psi[18];
    float *prof[24];
    float ****F[12];
    float **srcpos, **signals, **sec[3], **diff[3], **obsm, **recsrc;
    int **recpos;
    float *sig;
    ifos_geom geo;
    float ***rm[6] = {NULL}, ***taus = NULL, ***taup = NULL, ***tsa[3] = {NULL}, *eta = NULL;
    double t_begin = now_s(), t_loop = t_begin;

    NX = nx; NY = ny; NZ = nz; FW = fw; DX = dx; DY = dy; DZ = dz; DT = dt; NT = nt;
    TIME = (NT - 1) * DT;
    FDORDER = 4; FDCOEFF = fdcoeff; ABS_TYPE = 1; FREE_SURF = free_surf; L = g_taus ? 1 : 0; SEISMO = 1;
    POS[1] = POS[2] = POS[3] = 0; NPROCX = NPROCY = NPROCZ = 1; ALPHA = BETA = 0.0f;
    VPPML = vppml; FPML = fpml;

    /* ifos3d.c:288-348 extents (top rank) */
    vx = nr3(lo, NY + 2, -1, NX + 2, -1, NZ + 2); vy = nr3(lo, NY + 2, -1, NX + 2, -1, NZ + 2);
    vz = nr3(lo, NY + 2, -1, NX + 2, -1, NZ + 2); sxy = nr3(lo, NY + 2, -1, NX + 2, -1, NZ + 2);
    syz = nr3(lo, NY + 2, -1, NX + 2, -1, NZ + 2);
    sxx = nr3(-1, NY + 2, -1, NX + 2, -1, NZ + 2); syy = nr3(-1, NY + 2, -1, NX + 2, -1, NZ + 2);
    szz = nr3(-1, NY + 2, -1, NX + 2, -1, NZ + 2); sxz = nr3(-1, NY + 2, -1, NX + 2, -1, NZ + 2);
    rho = nr3(0, NY + 1, 0, NX + 1, 0, NZ + 1); pi = nr3(0, NY + 1, 0, NX + 1, 0, NZ + 1);
    u = nr3(0, NY + 1, 0, NX + 1, 0, NZ + 1);
    uipjp = nr3(1, NY, 1, NX, 1, NZ); ujpkp = nr3(1, NY, 1, NX, 1, NZ); uipkp = nr3(1, NY, 1, NX, 1, NZ);
    rjp = nr3(1, NY, 1, NX, 1, NZ); rkp = nr3(1, NY, 1, NX, 1, NZ); rip = nr3(1, NY, 1, NX, 1, NZ);
    g1 = nr3(1, NY, 1, NX, 1, NZ); g2 = nr3(1, NY, 1, NX, 1, NZ); g3 = nr3(1, NY, 1, NX, 1, NZ);
    absorb = nr3(1, NY, 1, NX, 1, NZ);
    for (n = 0; n < 18; n++) {
        int m = n % 9;
        psi[n] = m < 3 ? nr3(1, NY, 1, 2 * FW, 1, NZ) : m < 6 ? nr3(1, 2 * FW, 1, NX, 1, NZ) : nr3(1, NY, 1, NX, 1, 2 * FW);
    }
    for (n = 0; n < 24; n++) prof[n] = (float *)calloc((size_t)(2 * FW + 2), sizeof(float));
    for (n = 0; n < 12; n++) F[n] = nr4(1, nf, lo, NY + 2, -1, NX + 2, -1, NZ + 2);
    srcpos = nr2(1, 7, 1, 1);
    signals = nr2(1, 1, 1, NT);
    recpos = (int **)malloc(sizeof(int *) * 3) - 1;
    for (n = 1; n <= 3; n++) recpos[n] = (int *)calloc((size_t)ntr + 1, sizeof(int));
    recsrc = nr2(1, 7, 1, ntr);
    obsm = nr2(1, ntr, 1, NT);
    for (c = 0; c < 3; c++) {
        sec[c] = nr2(1, ntr, 1, NT);
        diff[c] = nr2(1, ntr, 1, NT);
    }

    /* model: host matcopy, then av_mat (ifos3d.c:681-683) */
    from_canon(rho, rho_c, 0, NY + 1, 0, NX + 1, 0, NZ + 1);
    from_canon(pi, pi_c, 0, NY + 1, 0, NX + 1, 0, NZ + 1);
    from_canon(u, u_c, 0, NY + 1, 0, NX + 1, 0, NZ + 1);
    matcopy1(rho); matcopy1(pi); matcopy1(u);
    if (L) {  /* ifos3d.c:445-457 */
        for (n = 0; n < 6; n++) rm[n] = nr3(1, NY, 1, NX, 1, NZ);
        taus = nr3(0, NY + 1, 0, NX + 1, 0, NZ + 1);
        taup = nr3(0, NY + 1, 0, NX + 1, 0, NZ + 1);
        for (n = 0; n < 3; n++) tsa[n] = nr3(1, NY, 1, NX, 1, NZ);
        eta = (float *)calloc(2, sizeof(float));
        eta[1] = g_eta;
        from_canon(taus, g_taus, 0, NY + 1, 0, NX + 1, 0, NZ + 1);
        from_canon(taup, g_taup, 0, NY + 1, 0, NX + 1, 0, NZ + 1);
        matcopy1(taus); matcopy1(taup);
    }
    av_mat(rho, pi, u, taus, taup, uipjp, ujpkp, uipkp, tsa[0], tsa[1], tsa[2], rjp, rkp, rip);
    CPML_ini_elastic(xb, yb, zb);
    /* per axis: K alpha a b K_half alpha_half a_half b_half (fd.h:57-62 order) */
    CPML_coeff(prof[0], prof[1], prof[2], prof[3], prof[4], prof[5], prof[6], prof[7],
               prof[8], prof[9], prof[10], prof[11], prof[12], prof[13], prof[14], prof[15],
               prof[16], prof[17], prof[18], prof[19], prof[20], prof[21], prof[22], prof[23]);

    /* source (splitsrc.c:44-50 rows) and wavelet (wavelet.c) */
    srcpos[1][1] = (float)src[0]; srcpos[2][1] = (float)src[1]; srcpos[3][1] = (float)src[2];
    srcpos[4][1] = src_f[0]; srcpos[5][1] = src_f[1]; srcpos[6][1] = src_f[2]; srcpos[7][1] = (float)src[3];
    ifos_geom_init(&geo, NX, NY, NZ, FW, DX, DY, DZ, DT, NT, FDCOEFF, FREE_SURF);
    sig = (float *)calloc((size_t)NT + 2, sizeof(float));
    ifos_wavelet(&geo, shape, src_f[0], src_f[1], src_f[2], sig);
    memcpy(&signals[1][1], &sig[1], sizeof(float) * NT);
    for (r = 1; r <= ntr; r++) {
        for (n = 1; n <= 3; n++) recpos[n][r] = rec_ijk[3 * (r - 1) + n - 1];
        for (n = 1; n <= 3; n++) recsrc[n][r] = (float)recpos[n][r];
    }
    ntast = obs ? ifos_ntast(finv[nf - 1], tast, DT) : 1;

#define PROF_V prof[0], prof[2], prof[3], prof[4], prof[6], prof[7], prof[8], prof[10], prof[11], prof[12], \
               prof[14], prof[15], prof[16], prof[18], prof[19], prof[20], prof[22], prof[23]
#define STRESS_UPDATE()                                                                                       \
    do {                                                                                                      \
        if (L) {                                                                                              \
            update_s(xb[0], xb[1], yb[0], yb[1], zb[0], zb[1], vx, vy, vz, sxx, syy, szz, sxy, syz, sxz, rm[0], \
                     rm[1], rm[2], rm[3], rm[4], rm[5], pi, u, uipjp, ujpkp, uipkp, taus, tsa[0], tsa[1],      \
                     tsa[2], taup, eta);                                                                      \
            update_s_CPML(xb[0], xb[1], yb[0], yb[1], zb[0], zb[1], vx, vy, vz, sxx, syy, szz, sxy, syz, sxz,   \
                          rm[0], rm[1], rm[2], rm[3], rm[4], rm[5], pi, u, uipjp, ujpkp, uipkp, taus, tsa[0],  \
                          tsa[1], tsa[2], taup, eta, PROF_V, psi[9], psi[10], psi[11], psi[12], psi[13],      \
                          psi[14], psi[15], psi[16], psi[17]);                                                \
        } else {                                                                                              \
            update_s_elastic(xb[0], xb[1], yb[0], yb[1], zb[0], zb[1], vx, vy, vz, sxx, syy, szz, sxy, syz,   \
                             sxz, pi, u, uipjp, ujpkp, uipkp);                                                \
            update_s_CPML_elastic(xb[0], xb[1], yb[0], yb[1], zb[0], zb[1], vx, vy, vz, sxx, syy, szz, sxy,   \
                                  syz, sxz, pi, u, uipjp, ujpkp, uipkp, PROF_V, psi[9], psi[10], psi[11],     \
                                  psi[12], psi[13], psi[14], psi[15], psi[16], psi[17]);                      \
        }                                                                                                     \
    } while (0)
    /* forward, ifos3d.c:738-814 */
    t_loop = now_s();
    zero_wavefield(NX, NY, NZ, vx, vy, vz, sxx, syy, szz, sxy, syz, sxz, rm[0], rm[1], rm[2], rm[3], rm[4], rm[5],
                   psi[0], psi[1], psi[2], psi[3], psi[4], psi[5], psi[6], psi[7], psi[8], psi[9], psi[10],
                   psi[11], psi[12], psi[13], psi[14], psi[15], psi[16], psi[17]);
    zero_invers(NX, NY, NZ, F[0], F[1], F[2], F[3], F[4], F[5], F[6], F[7], F[8], F[9], F[10], F[11], nf, 0);
    for (n = 1; n <= NT; n++) {
        if (isnan(vy[NY / 2][NX / 2][NZ / 2])) { rc = -1; goto done; }
        update_v(xb[0], xb[1], yb[0], yb[1], zb[0], zb[1], n, vx, vy, vz, sxx, syy, szz, sxy, syz, sxz, rho, rjp,
                 rkp, rip, srcpos, signals, signals, signals, 1, absorb, 0);
        update_v_CPML(xb[0], xb[1], yb[0], yb[1], zb[0], zb[1], n, vx, vy, vz, sxx, syy, szz, sxy, syz, sxz, rho,
                      rjp, rkp, rip, srcpos, signals, signals, signals, 1, absorb, 0, PROF_V, psi[0], psi[1],
                      psi[2], psi[3], psi[4], psi[5], psi[6], psi[7], psi[8]);
        STRESS_UPDATE();
        psource(n, sxx, syy, szz, srcpos, signals, 1);
        if (obs && n % ntast == 0)
            discfourier(1, NX, 1, NY, 1, NZ, n, vx, vy, vz, F[0], F[1], F[2], F[3], F[4], F[5], (float *)finv, nf,
                        ntast, 0, 0);
        if (FREE_SURF) surface_elastic(1, u, pi, sxx, syy, szz, sxy, syz, vx, vy, vz);
        seismo(n, ntr, recpos, sec[0], sec[1], sec[2], NULL, NULL, NULL, vx, vy, vz, sxx, syy, szz, pi, u);
    }
    for (c = 0; c < 3; c++)
        for (r = 1; r <= ntr; r++) memcpy(&syn[((size_t)c * ntr + r - 1) * (NT + 1) + 1], &sec[c][r][1], sizeof(float) * NT);
    if (!obs) goto done;

    /* residual, ifos3d.c:837-866 */
    *l2 = 0.0f;
    for (c = 0; c < 3; c++) {
        for (r = 1; r <= ntr; r++) memcpy(&obsm[r][1], &obs[((size_t)c * ntr + r - 1) * (NT + 1) + 1], sizeof(float) * NT);
        residual1(obsm, sec[c], diff[c], ntr, NT, l2);
    }
    for (n = 1; n <= 3; n++)
        for (r = 1; r <= ntr; r++) recsrc[n][r] = (float)recpos[n][r];

    /* adjoint, ifos3d.c:905-975 */
    zero_wavefield(NX, NY, NZ, vx, vy, vz, sxx, syy, szz, sxy, syz, sxz, rm[0], rm[1], rm[2], rm[3], rm[4], rm[5],
                   psi[0], psi[1], psi[2], psi[3], psi[4], psi[5], psi[6], psi[7], psi[8], psi[9], psi[10],
                   psi[11], psi[12], psi[13], psi[14], psi[15], psi[16], psi[17]);
    for (n = 1; n <= NT; n++) {
        if (isnan(vy[NY / 2][NX / 2][NZ / 2])) { rc = -1; goto done; }
        update_v(xb[0], xb[1], yb[0], yb[1], zb[0], zb[1], n, vx, vy, vz, sxx, syy, szz, sxy, syz, sxz, rho, rjp,
                 rkp, rip, recsrc, diff[0], diff[1], diff[2], ntr, absorb, 1);
        update_v_CPML(xb[0], xb[1], yb[0], yb[1], zb[0], zb[1], n, vx, vy, vz, sxx, syy, szz, sxy, syz, sxz, rho,
                      rjp, rkp, rip, recsrc, diff[0], diff[1], diff[2], ntr, absorb, 1, PROF_V, psi[0], psi[1],
                      psi[2], psi[3], psi[4], psi[5], psi[6], psi[7], psi[8]);
        STRESS_UPDATE();
        if (n % ntast == 0)
            discfourier(1, NX, 1, NY, 1, NZ, n, vx, vy, vz, F[6], F[7], F[8], F[9], F[10], F[11], (float *)finv, nf,
                        ntast, 0, 1);
        if (FREE_SURF) surface_elastic(1, u, pi, sxx, syy, szz, sxy, syz, vx, vy, vz);
    }
    from_canon(g1, grads, 1, NY, 1, NX, 1, NZ);
    from_canon(g2, grads + (size_t)(NY + 5) * (NX + 5) * (NZ + 5), 1, NY, 1, NX, 1, NZ);
    from_canon(g3, grads + 2 * (size_t)(NY + 5) * (NX + 5) * (NZ + 5), 1, NY, 1, NX, 1, NZ);
    gradient_F(NX, NY, NZ, F[0], F[1], F[2], F[3], F[4], F[5], F[6], F[7], F[8], F[9], F[10], F[11], g1, g2, g3,
               NT + 1, rho, pi, u, (float *)finv, nf, 1);
    dd_last_times[0] = now_s() - t_loop;
    dd_last_times[1] = t_loop - t_begin;
    to_canon(g1, grads, 1, NY, 1, NX, 1, NZ);
    to_canon(g2, grads + (size_t)(NY + 5) * (NX + 5) * (NZ + 5), 1, NY, 1, NX, 1, NZ);
    to_canon(g3, grads + 2 * (size_t)(NY + 5) * (NX + 5) * (NZ + 5), 1, NY, 1, NX, 1, NZ);

done:
    ifos_dropin_release();
    nr3_free(vx, lo, -1, -1); nr3_free(vy, lo, -1, -1); nr3_free(vz, lo, -1, -1);
    nr3_free(sxy, lo, -1, -1); nr3_free(syz, lo, -1, -1);
    nr3_free(sxx, -1, -1, -1); nr3_free(syy, -1, -1, -1); nr3_free(szz, -1, -1, -1); nr3_free(sxz, -1, -1, -1);
    nr3_free(rho, 0, 0, 0); nr3_free(pi, 0, 0, 0); nr3_free(u, 0, 0, 0);
    nr3_free(uipjp, 1, 1, 1); nr3_free(ujpkp, 1, 1, 1); nr3_free(uipkp, 1, 1, 1);
    nr3_free(rjp, 1, 1, 1); nr3_free(rkp, 1, 1, 1); nr3_free(rip, 1, 1, 1);
    nr3_free(g1, 1, 1, 1); nr3_free(g2, 1, 1, 1); nr3_free(g3, 1, 1, 1); nr3_free(absorb, 1, 1, 1);
    for (n = 0; n < 18; n++) nr3_free(psi[n], 1, 1, 1);
    for (n = 0; n < 24; n++) free(prof[n]);
    for (n = 0; n < 12; n++) nr4_free(F[n], 1, nf, lo, -1, -1);
    nr2_free(srcpos, 1, 1); nr2_free(signals, 1, 1); nr2_free(recsrc, 1, 1); nr2_free(obsm, 1, 1);
    for (c = 0; c < 3; c++) { nr2_free(sec[c], 1, 1); nr2_free(diff[c], 1, 1); }
    for (n = 1; n <= 3; n++) free(recpos[n]);
    free(recpos + 1);
    free(sig);
    if (L) {
        for (n = 0; n < 6; n++) nr3_free(rm[n], 1, 1, 1);
        nr3_free(taus, 0, 0, 0); nr3_free(taup, 0, 0, 0);
        for (n = 0; n < 3; n++) nr3_free(tsa[n], 1, 1, 1);
        free(eta);
    }
    (void)j;
    dd_last_times[2] = now_s() - t_begin;
    return rc;
}

/*
 * The Hessian wavefields and hess_F through the drop-in, in the order of
 * ifos3d.c:880-983: zero_invers with ntr_hess receiver-side sets, the
 * forward DFT (pshot 0, back 0) of nsteps velocity snapshots vf
 * [nsteps][3] (canonical), the receiver-side DFTs (pshot p = 1..ntr_hess,
 * back 1) of vb [ntr_hess][nsteps][3], then hess_F into hess [3]
 * (canonical, accumulated).  Each snapshot is a fresh tensor, as the drop-in
 * uploads a wavefield when it first sees it (on the reference's path the
 * velocities stay on the device between the updates).
 */
int dd_hess(int nx, int ny, int nz, int fw, float dx, float dy, float dz, float dt, int nt, int fdcoeff,
            const float *rho_c, const float *pi_c, const float *u_c, const float *vf, const float *vb, int nsteps,
            int ntr_hess, const float *finv, int nf, float *hess)
{
    const int lo = -2;
    const size_t cs = (size_t)(ny + 5) * (nx + 5) * (nz + 5);
    float ***rho, ***pi, ***u, ***h[3], ***v[3];
    float ****F[12];
    float ****snaps = (float ****)malloc(sizeof(float ***) * 3 * (size_t)(ntr_hess + 1) * nsteps);
    int n, p, s, c, ns_ = 0;

    NX = nx; NY = ny; NZ = nz; FW = fw; DX = dx; DY = dy; DZ = dz; DT = dt; NT = nt;
    TIME = (NT - 1) * DT;
    FDORDER = 4; FDCOEFF = fdcoeff; ABS_TYPE = 1; FREE_SURF = 0; L = 0; SEISMO = 1;
    POS[1] = POS[2] = POS[3] = 0; NPROCX = NPROCY = NPROCZ = 1; ALPHA = BETA = 0.0f;

    rho = nr3(0, NY + 1, 0, NX + 1, 0, NZ + 1); pi = nr3(0, NY + 1, 0, NX + 1, 0, NZ + 1);
    u = nr3(0, NY + 1, 0, NX + 1, 0, NZ + 1);
    from_canon(rho, rho_c, 0, NY + 1, 0, NX + 1, 0, NZ + 1);
    from_canon(pi, pi_c, 0, NY + 1, 0, NX + 1, 0, NZ + 1);
    from_canon(u, u_c, 0, NY + 1, 0, NX + 1, 0, NZ + 1);
    for (c = 0; c < 3; c++) h[c] = nr3(1, NY, 1, NX, 1, NZ);
    for (n = 0; n < 12; n++) F[n] = nr4(1, n < 6 ? nf : nf * (ntr_hess + 1), lo, NY + 2, -1, NX + 2, -1, NZ + 2);

    zero_invers(NX, NY, NZ, F[0], F[1], F[2], F[3], F[4], F[5], F[6], F[7], F[8], F[9], F[10], F[11], nf, ntr_hess);
    for (p = 0; p <= ntr_hess; p++)
        for (s = 0; s < nsteps; s++) {
            const float *src = p == 0 ? vf + (size_t)s * 3 * cs : vb + ((size_t)(p - 1) * nsteps + s) * 3 * cs;
            for (c = 0; c < 3; c++) {
                v[c] = snaps[ns_++] = nr3(lo, NY + 2, -1, NX + 2, -1, NZ + 2);
                from_canon(v[c], src + (size_t)c * cs, lo, NY + 2, -1, NX + 2, -1, NZ + 2);
            }
            if (p == 0)
                discfourier(1, NX, 1, NY, 1, NZ, s + 1, v[0], v[1], v[2], F[0], F[1], F[2], F[3], F[4], F[5],
                            (float *)finv, nf, 1, 0, 0);
            else
                discfourier(1, NX, 1, NY, 1, NZ, s + 1, v[0], v[1], v[2], F[6], F[7], F[8], F[9], F[10], F[11],
                            (float *)finv, nf, 1, p, 1);
            /* the tensors stay allocated until the release: the drop-in keys its mirrors by the pointer */
        }
    for (c = 0; c < 3; c++) from_canon(h[c], hess + (size_t)c * cs, 1, NY, 1, NX, 1, NZ);
    hess_F(NX, NY, NZ, F[0], F[1], F[2], F[3], F[4], F[5], F[6], F[7], F[8], F[9], F[10], F[11], h[0], h[1], h[2],
           NT, rho, pi, u, (float *)finv, nf, ntr_hess);
    for (c = 0; c < 3; c++) to_canon(h[c], hess + (size_t)c * cs, 1, NY, 1, NX, 1, NZ);
    ifos_dropin_release();
    for (n = 0; n < ns_; n++) nr3_free(snaps[n], lo, -1, -1);
    free(snaps);
    for (n = 0; n < 12; n++) nr4_free(F[n], 1, n < 6 ? nf : nf * (ntr_hess + 1), lo, -1, -1);
    for (c = 0; c < 3; c++) nr3_free(h[c], 1, 1, 1);
    nr3_free(rho, 0, 0, 0); nr3_free(pi, 0, 0, 0); nr3_free(u, 0, 0, 0);
    return 0;
}

/*
 * ifos3d.c:1005-1040 and :1138, iteration after iteration: the gradients
 * grads [niter][3] (canonical) go through lbfgs_savegrad / lbfgs (use_lbfgs 1)
 * or conjugate (use_lbfgs 0, cdf 1 on the first iteration), then modelupdate
 * with steps[it] updates the model model_c [3] (rho, pi, u, canonical).
 * Outputs: the model after the last update, the gradients as each iteration
 * leaves them (the direction, normalised when use_lbfgs is 0), beta [niter][3]
 * and bfgsscale [niter][bfgsnum].
 */
int dd_optim(int nx, int ny, int nz, int fw, float dx, float dt, int nt, const float *model_c, const float *grads_c,
             int niter, int use_lbfgs, int numpar, int bfgsnum, const float *steps, const float *weight, float vp0,
             float vs0, float rho0, float *model_out, float *dir_out, float *beta_out, float *scale_out)
{
    const size_t cs = (size_t)(ny + 5) * (nx + 5) * (nz + 5);
    float ***m[3], ***gr[3], ***prior[6], **bfgsmod, **bfgsgrad, *bfgsscale, beta[3] = {0.0f, 0.0f, 0.0f};
    int c, it, w;

    NX = nx; NY = ny; NZ = nz; FW = fw; DX = DY = DZ = dx; DT = dt; NT = nt;
    TIME = (NT - 1) * DT;
    FDORDER = 4; FDCOEFF = 2; ABS_TYPE = 1; FREE_SURF = 0; L = 0; SEISMO = 1;
    POS[1] = POS[2] = POS[3] = 0; NPROCX = NPROCY = NPROCZ = 1; ALPHA = BETA = 0.0f;
    LBFGS = use_lbfgs; NUMPAR = numpar; BFGSNUM = bfgsnum;
    VP0 = vp0; VS0 = vs0; RHO0 = rho0;
    for (c = 0; c < 3; c++) WEIGHT[c] = weight[c];

    for (c = 0; c < 3; c++) {
        m[c] = nr3(0, NY + 1, 0, NX + 1, 0, NZ + 1);
        from_canon(m[c], model_c + (size_t)c * cs, 0, NY + 1, 0, NX + 1, 0, NZ + 1);
        gr[c] = nr3(1, NY, 1, NX, 1, NZ);
    }
    for (c = 0; c < 6; c++) prior[c] = nr3(1, NY, 1, NX, 1, NZ);
    bfgsmod = nr2(1, BFGSNUM, 1, NUMPAR * NX * NY * NZ);
    bfgsgrad = nr2(1, BFGSNUM, 1, NUMPAR * NX * NY * NZ);
    bfgsscale = (float *)calloc((size_t)BFGSNUM + 1, sizeof(float));
    for (it = 0; it < niter; it++) {
        for (c = 0; c < 3; c++) from_canon(gr[c], grads_c + ((size_t)it * 3 + c) * cs, 1, NY, 1, NX, 1, NZ);
        if (LBFGS) {
            if (it > 0)
                lbfgs(gr[0], gr[1], gr[2], bfgsscale, bfgsmod, bfgsgrad, it + 1);
            else
                lbfgs_savegrad(gr[0], gr[1], gr[2], bfgsgrad);
        } else
            conjugate(NX, NY, NZ, gr[0], gr[1], gr[2], prior[0], prior[1], prior[2], prior[3], prior[4], prior[5],
                      beta, it + 1, it == 0);
        modelupdate(NX, NY, NZ, gr[0], gr[1], gr[2], m[0], m[1], m[2], bfgsmod, steps[it], beta, it + 1);
        for (c = 0; c < 3; c++) {
            to_canon(gr[c], dir_out + ((size_t)it * 3 + c) * cs, 1, NY, 1, NX, 1, NZ);
            beta_out[it * 3 + c] = beta[c];
        }
        for (w = 1; w <= BFGSNUM; w++) scale_out[it * BFGSNUM + w - 1] = bfgsscale[w];
    }
    for (c = 0; c < 3; c++) to_canon(m[c], model_out + (size_t)c * cs, 0, NY + 1, 0, NX + 1, 0, NZ + 1);
    ifos_dropin_release();
    for (c = 0; c < 3; c++) {
        nr3_free(m[c], 0, 0, 0);
        nr3_free(gr[c], 1, 1, 1);
    }
    for (c = 0; c < 6; c++) nr3_free(prior[c], 1, 1, 1);
    nr2_free(bfgsmod, 1, 1);
    nr2_free(bfgsgrad, 1, 1);
    free(bfgsscale);
    return 0;
}
