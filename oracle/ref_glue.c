/*
 * ref_glue.c -- harness around the compiled IFOS3D reference routines
 * (TEST INFRASTRUCTURE, built only by oracle/Makefile.ref into oracle/_ref/).
 *
 * It owns the reference globals (by including the reference's globvar.h, as
 * ifos3d.c:26 does), allocates the reference's NR tensors with the
 * reference's own f3tensor/f4tensor over the extents of ifos3d.c:288-348,
 * converts them to and from the canonical flat layout of ifos_oracle.h, and
 * calls the reference routines one by one or in the shot-loop order of
 * ifos3d.c:712-975.  Single rank: POS = 0, NPROC* = 1, neighbours
 * MPI_PROC_NULL (initproc.c:73-78).
 */
#include "fd.h"
#include "globvar.h"

/* field ids for rg_set / rg_get */
enum {
    RG_VX, RG_VY, RG_VZ, RG_SXX, RG_SYY, RG_SZZ, RG_SXY, RG_SYZ, RG_SXZ, /* 0-8 */
    RG_RHO, RG_PI, RG_U,                                                /* 9-11 */
    RG_UIPJP, RG_UJPKP, RG_UIPKP, RG_RJP, RG_RKP, RG_RIP,               /* 12-17 */
    RG_GRAD1, RG_GRAD2, RG_GRAD3,                                       /* 18-20 */
    RG_RXX, RG_RYY, RG_RZZ, RG_RXY, RG_RYZ, RG_RXZ,                     /* 21-26 */
    RG_TAUS, RG_TAUP, RG_TAUSIPJP, RG_TAUSJPKP, RG_TAUSIPKP,            /* 27-31 */
    RG_NFIELDS
};

static struct {
    int ready;
    float ***f[RG_NFIELDS];
    int lo[RG_NFIELDS][3], hi[RG_NFIELDS][3]; /* j, i, k bounds */
    /* psi: 0-8 stress-derivative memory (psi_sxx_x ... psi_szz_z),
     *      9-17 velocity-derivative memory (psi_vxx ... psi_vzz) */
    float ***psi[18];
    float *prof[24];          /* K a b K_half a_half b_half alpha alpha_half per axis */
    float ****F[12];          /* Ffvx Ffvy Ffvz Ffivx Ffivy Ffivz Fbvx ... Fbivz */
    int nfmax;
    int xb[2], yb[2], zb[2];
    float **srcpos;           /* [1..7][1..nsrc] */
    float **signals;          /* [1..nsrc][1..NT] */
    float **adjx, **adjy, **adjz; /* [1..ntr][1..NT] */
    float **recsrc;           /* [1..7][1..ntr] receiver positions as sources */
    int nsrc, ntr;
    int **recpos;             /* [1..3][1..ntr] */
    float **secx, **secy, **secz; /* [1..ntr][1..NT] */
    float ***absorb;
    float *eta;               /* [1..1] relaxation parameter of the mechanism */
    float *finv;
    int nf;
} S;

static void alloc3(int id, int jl, int jh, int il, int ih, int kl, int kh)
{
    S.f[id] = f3tensor(jl, jh, il, ih, kl, kh);
    S.lo[id][0] = jl; S.hi[id][0] = jh;
    S.lo[id][1] = il; S.hi[id][1] = ih;
    S.lo[id][2] = kl; S.hi[id][2] = kh;
}

static size_t cidx(int j, int i, int k)
{
    return ((size_t)(j + 2) * (size_t)(NX + 5) + (size_t)(i + 2)) * (size_t)(NZ + 5) + (size_t)(k + 2);
}

/* set the reference globals and allocate every array of one shot */
int rg_setup(int nx, int ny, int nz, float dx, float dy, float dz, float dt,
             int nt, int fw, int fdcoeff, int free_surf, float vppml,
             float fpml, int nfmax, int nsrc, int ntr)
{
    int flag = 0, n, lo, l = 1;
    static void *bsend_buf = NULL;
    static int bsend_size = 0;
    int need;
    MPI_Initialized(&flag);
    if (!flag) MPI_Init(NULL, NULL);
    /* matcopy() sends with MPI_Bsend even to MPI_PROC_NULL, which needs an
     * attached buffer (ifos3d.c:266-270 attaches one) */
    need = 2 * (int)(((size_t)(nx + 2) * (ny + 2) + (size_t)(ny + 2) * (nz + 2) +
                      (size_t)(nx + 2) * (nz + 2)) * 5 * sizeof(float)) + 16 * MPI_BSEND_OVERHEAD;
    if (need > bsend_size) {
        if (bsend_buf) {
            void *old; int osz;
            MPI_Buffer_detach(&old, &osz);
            free(old);
        }
        bsend_buf = malloc((size_t)need);
        bsend_size = need;
        MPI_Buffer_attach(bsend_buf, need);
    }

    NX = nx; NY = ny; NZ = nz; NXG = nx; NYG = ny; NZG = nz;
    DX = dx; DY = dy; DZ = dz; DT = dt;
    NT = nt; TIME = (NT - 1) * DT;          /* ifos3d.c:176-177 */
    FW = fw; FDORDER = 4; FDCOEFF = fdcoeff; FREE_SURF = free_surf;
    ABS_TYPE = 1; L = 0; VPPML = vppml; FPML = fpml;
    SEISMO = 1; NDT = 1; NDTSHIFT = 0; METHOD = 1; BOUNDARY = 0;
    ALPHA = 0.0f; BETA = 0.0f; VERBOSE = 0; MYID = 0;
    NPROC = NP = 1; NPROCX = NPROCY = NPROCZ = 1;
    IENDX = nx; IENDY = ny; IENDZ = nz;
    POS[1] = POS[2] = POS[3] = 0;
    for (n = 1; n <= 6; n++) INDEX[n] = MPI_PROC_NULL;
    FP = stderr;

    /* wavefields, ifos3d.c:288-348 (top rank: vx vy vz sxy syz start at 0-FDORDER/2) */
    lo = 0 - FDORDER / 2;
    alloc3(RG_VX, lo, NY + l * FDORDER / 2, 1 - l * FDORDER / 2, NX + l * FDORDER / 2, 1 - l * FDORDER / 2, NZ + l * FDORDER / 2);
    alloc3(RG_VY, lo, NY + l * FDORDER / 2, 1 - l * FDORDER / 2, NX + l * FDORDER / 2, 1 - l * FDORDER / 2, NZ + l * FDORDER / 2);
    alloc3(RG_VZ, lo, NY + l * FDORDER / 2, 1 - l * FDORDER / 2, NX + l * FDORDER / 2, 1 - l * FDORDER / 2, NZ + l * FDORDER / 2);
    alloc3(RG_SXY, lo, NY + l * FDORDER / 2, 1 - l * FDORDER / 2, NX + l * FDORDER / 2, 1 - l * FDORDER / 2, NZ + l * FDORDER / 2);
    alloc3(RG_SYZ, lo, NY + l * FDORDER / 2, 1 - l * FDORDER / 2, NX + l * FDORDER / 2, 1 - l * FDORDER / 2, NZ + l * FDORDER / 2);
    alloc3(RG_SXZ, 1 - l * FDORDER / 2, NY + l * FDORDER / 2, 1 - l * FDORDER / 2, NX + l * FDORDER / 2, 1 - l * FDORDER / 2, NZ + l * FDORDER / 2);
    alloc3(RG_SXX, 1 - l * FDORDER / 2, NY + l * FDORDER / 2, 1 - l * FDORDER / 2, NX + l * FDORDER / 2, 1 - l * FDORDER / 2, NZ + l * FDORDER / 2);
    alloc3(RG_SYY, 1 - l * FDORDER / 2, NY + l * FDORDER / 2, 1 - l * FDORDER / 2, NX + l * FDORDER / 2, 1 - l * FDORDER / 2, NZ + l * FDORDER / 2);
    alloc3(RG_SZZ, 1 - l * FDORDER / 2, NY + l * FDORDER / 2, 1 - l * FDORDER / 2, NX + l * FDORDER / 2, 1 - l * FDORDER / 2, NZ + l * FDORDER / 2);
    alloc3(RG_RHO, 0, NY + 1, 0, NX + 1, 0, NZ + 1);
    alloc3(RG_PI, 0, NY + 1, 0, NX + 1, 0, NZ + 1);
    alloc3(RG_U, 0, NY + 1, 0, NX + 1, 0, NZ + 1);
    for (n = RG_UIPJP; n <= RG_GRAD3; n++) alloc3(n, 1, NY, 1, NX, 1, NZ);
    /* viscoelastic arrays, ifos3d.c:445-457 */
    for (n = RG_RXX; n <= RG_RXZ; n++) alloc3(n, 1, NY, 1, NX, 1, NZ);
    alloc3(RG_TAUS, 0, NY + 1, 0, NX + 1, 0, NZ + 1);
    alloc3(RG_TAUP, 0, NY + 1, 0, NX + 1, 0, NZ + 1);
    for (n = RG_TAUSIPJP; n <= RG_TAUSIPKP; n++) alloc3(n, 1, NY, 1, NX, 1, NZ);
    S.eta = vector(1, 1);
    S.absorb = f3tensor(1, NY, 1, NX, 1, NZ);

    /* CPML memory, ifos3d.c:421-440 */
    for (n = 0; n < 3; n++) S.psi[n] = f3tensor(1, NY, 1, 2 * FW, 1, NZ);
    for (n = 3; n < 6; n++) S.psi[n] = f3tensor(1, 2 * FW, 1, NX, 1, NZ);
    for (n = 6; n < 9; n++) S.psi[n] = f3tensor(1, NY, 1, NX, 1, 2 * FW);
    for (n = 9; n < 12; n++) S.psi[n] = f3tensor(1, NY, 1, 2 * FW, 1, NZ);
    for (n = 12; n < 15; n++) S.psi[n] = f3tensor(1, 2 * FW, 1, NX, 1, NZ);
    for (n = 15; n < 18; n++) S.psi[n] = f3tensor(1, NY, 1, NX, 1, 2 * FW);
    for (n = 0; n < 24; n++) S.prof[n] = vector(1, 2 * FW);

    S.nfmax = nfmax;
    for (n = 0; n < 12; n++)
        S.F[n] = f4tensor(1, nfmax, lo, NY + l * FDORDER / 2, 1 - l * FDORDER / 2, NX + l * FDORDER / 2,
                          1 - l * FDORDER / 2, NZ + l * FDORDER / 2);
    S.finv = vector(0, nfmax - 1);

    S.nsrc = nsrc;
    S.ntr = ntr;
    S.srcpos = fmatrix(1, 7, 1, nsrc > 0 ? nsrc : 1);
    S.signals = fmatrix(1, nsrc > 0 ? nsrc : 1, 1, NT);
    S.recpos = imatrix(1, 3, 1, ntr > 0 ? ntr : 1);
    S.recsrc = fmatrix(1, 7, 1, ntr > 0 ? ntr : 1);
    S.adjx = fmatrix(1, ntr > 0 ? ntr : 1, 1, NT);
    S.adjy = fmatrix(1, ntr > 0 ? ntr : 1, 1, NT);
    S.adjz = fmatrix(1, ntr > 0 ? ntr : 1, 1, NT);
    S.secx = fmatrix(1, ntr > 0 ? ntr : 1, 1, NT);
    S.secy = fmatrix(1, ntr > 0 ? ntr : 1, 1, NT);
    S.secz = fmatrix(1, ntr > 0 ? ntr : 1, 1, NT);

    CPML_ini_elastic(S.xb, S.yb, S.zb);
    CPML_coeff(S.prof[0], S.prof[6], S.prof[1], S.prof[2], S.prof[3], S.prof[7], S.prof[4], S.prof[5],
               S.prof[8], S.prof[14], S.prof[9], S.prof[10], S.prof[11], S.prof[15], S.prof[12], S.prof[13],
               S.prof[16], S.prof[22], S.prof[17], S.prof[18], S.prof[19], S.prof[23], S.prof[20], S.prof[21]);
    S.ready = 1;
    return 0;
}

/* canonical <-> NR over the NR array's extents */
void rg_set(int id, const float *canon)
{
    int j, i, k;
    for (j = S.lo[id][0]; j <= S.hi[id][0]; j++)
        for (i = S.lo[id][1]; i <= S.hi[id][1]; i++)
            for (k = S.lo[id][2]; k <= S.hi[id][2]; k++) S.f[id][j][i][k] = canon[cidx(j, i, k)];
}

void rg_get(int id, float *canon)
{
    int j, i, k;
    for (j = S.lo[id][0]; j <= S.hi[id][0]; j++)
        for (i = S.lo[id][1]; i <= S.hi[id][1]; i++)
            for (k = S.lo[id][2]; k <= S.hi[id][2]; k++) canon[cidx(j, i, k)] = S.f[id][j][i][k];
}

/* psi arrays as compact 0-based blocks (same shapes as the oracle's) */
static void psi_dims(int n, int *a, int *b, int *c)
{
    int m = n % 9;
    if (m < 3) { *a = NY; *b = 2 * FW; *c = NZ; }
    else if (m < 6) { *a = 2 * FW; *b = NX; *c = NZ; }
    else { *a = NY; *b = NX; *c = 2 * FW; }
}

void rg_psi_set(int n, const float *p)
{
    int a, b, c, x, y, z;
    psi_dims(n, &a, &b, &c);
    for (x = 1; x <= a; x++)
        for (y = 1; y <= b; y++)
            for (z = 1; z <= c; z++) S.psi[n][x][y][z] = p[((size_t)(x - 1) * b + (y - 1)) * c + (z - 1)];
}

void rg_psi_get(int n, float *p)
{
    int a, b, c, x, y, z;
    psi_dims(n, &a, &b, &c);
    for (x = 1; x <= a; x++)
        for (y = 1; y <= b; y++)
            for (z = 1; z <= c; z++) p[((size_t)(x - 1) * b + (y - 1)) * c + (z - 1)] = S.psi[n][x][y][z];
}

/* profile n (0..23, order K a b K_h a_h b_h alpha alpha_h per axis) into
 * out[0..2fw] (out[0] = 0) */
void rg_prof_get(int n, float *out)
{
    int h;
    out[0] = 0.0f;
    for (h = 1; h <= 2 * FW; h++) out[h] = S.prof[n][h];
}

void rg_box(int *box)
{
    box[0] = S.xb[0]; box[1] = S.xb[1];
    box[2] = S.yb[0]; box[3] = S.yb[1];
    box[4] = S.zb[0]; box[5] = S.zb[1];
}

/* frequency fields: component n (0..11), frequency m (1-based) */
void rg_F_set(int n, int m, const float *canon)
{
    int j, i, k;
    for (j = -2; j <= NY + 2; j++)
        for (i = -1; i <= NX + 2; i++)
            for (k = -1; k <= NZ + 2; k++) S.F[n][m][j][i][k] = canon[cidx(j, i, k)];
}

void rg_F_get(int n, int m, float *canon)
{
    int j, i, k;
    for (j = -2; j <= NY + 2; j++)
        for (i = -1; i <= NX + 2; i++)
            for (k = -1; k <= NZ + 2; k++) canon[cidx(j, i, k)] = S.F[n][m][j][i][k];
}

void rg_set_freqs(const float *finv, int nf)
{
    int l;
    S.nf = nf;
    for (l = 0; l < nf; l++) S.finv[l] = finv[l];
}

/* one source: grid indices, tshift, fc, amp, type (srcpos_loc rows 1-7) */
void rg_set_source(int s, int i, int j, int k, float tshift, float fc, float amp, int type)
{
    S.srcpos[1][s + 1] = (float)i;
    S.srcpos[2][s + 1] = (float)j;
    S.srcpos[3][s + 1] = (float)k;
    S.srcpos[4][s + 1] = tshift;
    S.srcpos[5][s + 1] = fc;
    S.srcpos[6][s + 1] = amp;
    S.srcpos[7][s + 1] = (float)type;
}

void rg_make_wavelet(int shape)
{
    wavelet(S.srcpos, S.nsrc, shape, S.signals);
}

/* signals as [nsrc][nt+2] indexed by nt */
void rg_get_signals(float *out)
{
    int s, n;
    for (s = 1; s <= S.nsrc; s++)
        for (n = 1; n <= NT; n++) out[(size_t)(s - 1) * (NT + 2) + n] = S.signals[s][n];
}

void rg_set_signals(const float *in)
{
    int s, n;
    for (s = 1; s <= S.nsrc; s++)
        for (n = 1; n <= NT; n++) S.signals[s][n] = in[(size_t)(s - 1) * (NT + 2) + n];
}

void rg_set_receivers(const int *ijk)
{
    int r;
    for (r = 1; r <= S.ntr; r++) {
        S.recpos[1][r] = ijk[3 * (r - 1) + 0];
        S.recpos[2][r] = ijk[3 * (r - 1) + 1];
        S.recpos[3][r] = ijk[3 * (r - 1) + 2];
        S.recsrc[1][r] = (float)S.recpos[1][r];
        S.recsrc[2][r] = (float)S.recpos[2][r];
        S.recsrc[3][r] = (float)S.recpos[3][r];
    }
}

/* adjoint traces [ntr][nt+2] indexed by nt */
void rg_set_adjoint(const float *ax, const float *ay, const float *az)
{
    int r, n;
    for (r = 1; r <= S.ntr; r++)
        for (n = 1; n <= NT; n++) {
            S.adjx[r][n] = ax[(size_t)(r - 1) * (NT + 2) + n];
            S.adjy[r][n] = ay[(size_t)(r - 1) * (NT + 2) + n];
            S.adjz[r][n] = az[(size_t)(r - 1) * (NT + 2) + n];
        }
}

/* sections as [ntr][nt+1] indexed by sample */
void rg_get_sections(float *sx, float *sy, float *sz)
{
    int r, n;
    for (r = 1; r <= S.ntr; r++)
        for (n = 1; n <= NT; n++) {
            sx[(size_t)(r - 1) * (NT + 1) + n] = S.secx[r][n];
            sy[(size_t)(r - 1) * (NT + 1) + n] = S.secy[r][n];
            sz[(size_t)(r - 1) * (NT + 1) + n] = S.secz[r][n];
        }
}

void rg_set_sections(const float *sx, const float *sy, const float *sz)
{
    int r, n;
    for (r = 1; r <= S.ntr; r++)
        for (n = 1; n <= NT; n++) {
            S.secx[r][n] = sx[(size_t)(r - 1) * (NT + 1) + n];
            S.secy[r][n] = sy[(size_t)(r - 1) * (NT + 1) + n];
            S.secz[r][n] = sz[(size_t)(r - 1) * (NT + 1) + n];
        }
}

/* adjoint traces as [ntr][nt+2] by step */
void rg_get_adjoint(float *ax, float *ay, float *az)
{
    int r, n;
    for (r = 1; r <= S.ntr; r++)
        for (n = 1; n <= NT; n++) {
            ax[(size_t)(r - 1) * (NT + 2) + n] = S.adjx[r][n];
            ay[(size_t)(r - 1) * (NT + 2) + n] = S.adjy[r][n];
            az[(size_t)(r - 1) * (NT + 2) + n] = S.adjz[r][n];
        }
}

/* --- single reference calls ------------------------------------------ */
#define W S.f
void rg_update_v(int nt, int back)
{
    if (back == 0)
        update_v(S.xb[0], S.xb[1], S.yb[0], S.yb[1], S.zb[0], S.zb[1], nt, W[RG_VX], W[RG_VY], W[RG_VZ],
                 W[RG_SXX], W[RG_SYY], W[RG_SZZ], W[RG_SXY], W[RG_SYZ], W[RG_SXZ], W[RG_RHO], W[RG_RJP],
                 W[RG_RKP], W[RG_RIP], S.srcpos, S.signals, S.signals, S.signals, S.nsrc, S.absorb, 0);
    else
        update_v(S.xb[0], S.xb[1], S.yb[0], S.yb[1], S.zb[0], S.zb[1], nt, W[RG_VX], W[RG_VY], W[RG_VZ],
                 W[RG_SXX], W[RG_SYY], W[RG_SZZ], W[RG_SXY], W[RG_SYZ], W[RG_SXZ], W[RG_RHO], W[RG_RJP],
                 W[RG_RKP], W[RG_RIP], S.recsrc, S.adjx, S.adjy, S.adjz, S.ntr, S.absorb, 1);
}

void rg_update_v_cpml(int nt, int back)
{
    float **sp = back ? S.recsrc : S.srcpos;
    float **s0 = back ? S.adjx : S.signals, **s1 = back ? S.adjy : S.signals, **s2 = back ? S.adjz : S.signals;
    int ns = back ? S.ntr : S.nsrc;
    float **P = S.prof;
    update_v_CPML(S.xb[0], S.xb[1], S.yb[0], S.yb[1], S.zb[0], S.zb[1], nt, W[RG_VX], W[RG_VY], W[RG_VZ],
                  W[RG_SXX], W[RG_SYY], W[RG_SZZ], W[RG_SXY], W[RG_SYZ], W[RG_SXZ], W[RG_RHO], W[RG_RJP],
                  W[RG_RKP], W[RG_RIP], sp, s0, s1, s2, ns, S.absorb, back,
                  P[0], P[1], P[2], P[3], P[4], P[5], P[8], P[9], P[10], P[11], P[12], P[13],
                  P[16], P[17], P[18], P[19], P[20], P[21],
                  S.psi[0], S.psi[1], S.psi[2], S.psi[3], S.psi[4], S.psi[5], S.psi[6], S.psi[7], S.psi[8]);
}

void rg_update_s(void)
{
    update_s_elastic(S.xb[0], S.xb[1], S.yb[0], S.yb[1], S.zb[0], S.zb[1], W[RG_VX], W[RG_VY], W[RG_VZ],
                     W[RG_SXX], W[RG_SYY], W[RG_SZZ], W[RG_SXY], W[RG_SYZ], W[RG_SXZ], W[RG_PI], W[RG_U],
                     W[RG_UIPJP], W[RG_UJPKP], W[RG_UIPKP]);
}

void rg_update_s_cpml(void)
{
    float **P = S.prof;
    update_s_CPML_elastic(S.xb[0], S.xb[1], S.yb[0], S.yb[1], S.zb[0], S.zb[1], W[RG_VX], W[RG_VY], W[RG_VZ],
                          W[RG_SXX], W[RG_SYY], W[RG_SZZ], W[RG_SXY], W[RG_SYZ], W[RG_SXZ], W[RG_PI], W[RG_U],
                          W[RG_UIPJP], W[RG_UJPKP], W[RG_UIPKP],
                          P[0], P[1], P[2], P[3], P[4], P[5], P[8], P[9], P[10], P[11], P[12], P[13],
                          P[16], P[17], P[18], P[19], P[20], P[21],
                          S.psi[9], S.psi[10], S.psi[11], S.psi[12], S.psi[13], S.psi[14], S.psi[15],
                          S.psi[16], S.psi[17]);
}

void rg_psource(int nt)
{
    psource(nt, W[RG_SXX], W[RG_SYY], W[RG_SZZ], S.srcpos, S.signals, S.nsrc);
}

void rg_surface(void)
{
    surface_elastic(1, W[RG_U], W[RG_PI], W[RG_SXX], W[RG_SYY], W[RG_SZZ], W[RG_SXY], W[RG_SYZ],
                    W[RG_VX], W[RG_VY], W[RG_VZ]);
}

void rg_seismo(int ins)
{
    seismo(ins, S.ntr, S.recpos, S.secx, S.secy, S.secz, NULL, NULL, NULL, W[RG_VX], W[RG_VY], W[RG_VZ],
           W[RG_SXX], W[RG_SYY], W[RG_SZZ], W[RG_PI], W[RG_U]);
}

void rg_discfourier(int nt, int ntast, int back)
{
    int o = back ? 6 : 0;
    discfourier(1, NX, 1, NY, 1, NZ, nt, W[RG_VX], W[RG_VY], W[RG_VZ], S.F[o + 0], S.F[o + 1], S.F[o + 2],
                S.F[o + 3], S.F[o + 4], S.F[o + 5], S.finv, S.nf, ntast, 0, back);
}

void rg_gradient_F(void)
{
    gradient_F(NX, NY, NZ, S.F[0], S.F[1], S.F[2], S.F[3], S.F[4], S.F[5], S.F[6], S.F[7], S.F[8], S.F[9],
               S.F[10], S.F[11], W[RG_GRAD1], W[RG_GRAD2], W[RG_GRAD3], NT + 1, W[RG_RHO], W[RG_PI], W[RG_U],
               S.finv, S.nf, 1);
}

/* hess_F over the F arrays (receiver n, frequency l at index n*nf + l),
 * accumulated into grad1..3 */
void rg_hess_F(int ntr_hess)
{
    hess_F(NX, NY, NZ, S.F[0], S.F[1], S.F[2], S.F[3], S.F[4], S.F[5], S.F[6], S.F[7], S.F[8], S.F[9],
           S.F[10], S.F[11], W[RG_GRAD1], W[RG_GRAD2], W[RG_GRAD3], NT + 1, W[RG_RHO], W[RG_PI], W[RG_U],
           S.finv, S.nf, ntr_hess);
}

void rg_matcopy(void)
{
    matcopy(W[RG_RHO], W[RG_PI], W[RG_U], W[RG_TAUS], W[RG_TAUP]);
}

void rg_av_mat(void)
{
    av_mat(W[RG_RHO], W[RG_PI], W[RG_U], W[RG_TAUS], W[RG_TAUP], W[RG_UIPJP], W[RG_UJPKP], W[RG_UIPKP],
           W[RG_TAUSIPJP], W[RG_TAUSJPKP], W[RG_TAUSIPKP], W[RG_RJP], W[RG_RKP], W[RG_RIP]);
}

/* the global L (number of relaxation mechanisms; the reference supports
 * 0 and 1) and eta[1] */
void rg_set_visco(int l, float eta)
{
    L = l;
    S.eta[1] = eta;
}

void rg_update_s_visco(void)
{
    update_s(S.xb[0], S.xb[1], S.yb[0], S.yb[1], S.zb[0], S.zb[1], W[RG_VX], W[RG_VY], W[RG_VZ], W[RG_SXX],
             W[RG_SYY], W[RG_SZZ], W[RG_SXY], W[RG_SYZ], W[RG_SXZ], W[RG_RXX], W[RG_RYY], W[RG_RZZ], W[RG_RXY],
             W[RG_RYZ], W[RG_RXZ], W[RG_PI], W[RG_U], W[RG_UIPJP], W[RG_UJPKP], W[RG_UIPKP], W[RG_TAUS],
             W[RG_TAUSIPJP], W[RG_TAUSJPKP], W[RG_TAUSIPKP], W[RG_TAUP], S.eta);
}

void rg_update_s_cpml_visco(void)
{
    float **P = S.prof;
    update_s_CPML(S.xb[0], S.xb[1], S.yb[0], S.yb[1], S.zb[0], S.zb[1], W[RG_VX], W[RG_VY], W[RG_VZ],
                  W[RG_SXX], W[RG_SYY], W[RG_SZZ], W[RG_SXY], W[RG_SYZ], W[RG_SXZ], W[RG_RXX], W[RG_RYY],
                  W[RG_RZZ], W[RG_RXY], W[RG_RYZ], W[RG_RXZ], W[RG_PI], W[RG_U], W[RG_UIPJP], W[RG_UJPKP],
                  W[RG_UIPKP], W[RG_TAUS], W[RG_TAUSIPJP], W[RG_TAUSJPKP], W[RG_TAUSIPKP], W[RG_TAUP], S.eta,
                  P[0], P[1], P[2], P[3], P[4], P[5], P[8], P[9], P[10], P[11], P[12], P[13],
                  P[16], P[17], P[18], P[19], P[20], P[21],
                  S.psi[9], S.psi[10], S.psi[11], S.psi[12], S.psi[13], S.psi[14], S.psi[15],
                  S.psi[16], S.psi[17]);
}

void rg_zero_wavefield(void)
{
    zero_wavefield(NX, NY, NZ, W[RG_VX], W[RG_VY], W[RG_VZ], W[RG_SXX], W[RG_SYY], W[RG_SZZ], W[RG_SXY],
                   W[RG_SYZ], W[RG_SXZ], W[RG_RXX], W[RG_RYY], W[RG_RZZ], W[RG_RXY], W[RG_RYZ], W[RG_RXZ], S.psi[0], S.psi[1], S.psi[2],
                   S.psi[3], S.psi[4], S.psi[5], S.psi[6], S.psi[7], S.psi[8], S.psi[9], S.psi[10], S.psi[11],
                   S.psi[12], S.psi[13], S.psi[14], S.psi[15], S.psi[16], S.psi[17]);
}

void rg_zero_invers(void)
{
    zero_invers(NX, NY, NZ, S.F[0], S.F[1], S.F[2], S.F[3], S.F[4], S.F[5], S.F[6], S.F[7], S.F[8], S.F[9],
                S.F[10], S.F[11], S.nfmax, 0);
}

void rg_zero_grad(void)
{
    zero_grad(NX, NY, NZ, W[RG_GRAD1], W[RG_GRAD2], W[RG_GRAD3]);
}

/* residual over ntr traces: obs/syn [ntr][nt+1] by sample; writes the
 * adjoint traces [ntr][nt+2] by step and returns the running L2 */
float rg_residual(const float *obs_x, const float *obs_y, const float *obs_z, float l2)
{
    float **obs = fmatrix(1, S.ntr, 1, NT), **obsf = fmatrix(1, S.ntr, 1, 1);
    float l2f = 0.0f;
    int r, n, c;
    const float *o[3] = {obs_x, obs_y, obs_z};
    float **sec[3] = {S.secx, S.secy, S.secz}, **adj[3] = {S.adjx, S.adjy, S.adjz};
    for (c = 0; c < 3; c++) {
        for (r = 1; r <= S.ntr; r++)
            for (n = 1; n <= NT; n++) obs[r][n] = o[c][(size_t)(r - 1) * (NT + 1) + n];
        residual(obs, obsf, sec[c], adj[c], S.ntr, NT, &l2, &l2f);
    }
    free_matrix(obs, 1, S.ntr, 1, NT);
    free_matrix(obsf, 1, S.ntr, 1, 1);
    return l2;
}

/* one gradient shot in the order of ifos3d.c:712-975 (forward, residual,
 * adjoint, gradient_F).  The model (rho, pi, u) must already be set; the
 * source is source 0.  Returns the L2 misfit, or -1 on the NaN probe. */
float rg_run_shot(int shape, int tast, const float *obs_x, const float *obs_y, const float *obs_z,
                  int want_adjoint)
{
    int nt, ntast;
    float dummy, l2 = 0.0f;
    rg_matcopy();
    rg_av_mat();
    rg_make_wavelet(shape);
    rg_zero_wavefield();
    rg_zero_invers();
    dummy = (1 / (S.finv[S.nf - 1] * tast * DT));
    ntast = (int)dummy;
    if (!ntast) ntast = 1;
    for (nt = 1; nt <= NT; nt++) {
        if (isnan(W[RG_VY][NY / 2][NX / 2][NZ / 2])) return -1.0f;
        rg_update_v(nt, 0);
        rg_update_v_cpml(nt, 0);
        rg_update_s();
        rg_update_s_cpml();
        rg_psource(nt);
        if (want_adjoint && nt % ntast == 0) rg_discfourier(nt, ntast, 0);
        if (FREE_SURF) rg_surface();
        rg_seismo(nt);
    }
    if (!want_adjoint) return 0.0f;
    if (obs_x) l2 = rg_residual(obs_x, obs_y, obs_z, 0.0f);
    rg_zero_wavefield();
    for (nt = 1; nt <= NT; nt++) {
        if (isnan(W[RG_VY][NY / 2][NX / 2][NZ / 2])) return -1.0f;
        rg_update_v(nt, 1);
        rg_update_v_cpml(nt, 1);
        rg_update_s();
        rg_update_s_cpml();
        if (nt % ntast == 0) rg_discfourier(nt, ntast, 1);
        if (FREE_SURF) rg_surface();
    }
    rg_gradient_F();
    return l2;
}

/* outseis.c with SEIS_FORMAT 1 into path: sec [ntr][ns+1] by sample,
 * rec_xyz [ntr][3] grid indices of trace t's receiver, rec_id [ntr] its
 * global number (recpos_loc[4][t]); src_xyz the one source (nsrc = 1, as
 * saveseis.c:38-39 passes it) or NULL for nsrc = 0 */
void rg_outseis(const char *path, int comp, const float *sec, int ntr, int ns, const int *rec_xyz,
                const int *rec_id, const float *src_xyz, float dx, float dy, float dz, float dt, int ndt)
{
    int t, j, nrec = 1;
    float **section, **srcpos;
    int **recpos, **recpos_loc;
    for (t = 0; t < ntr; t++)
        if (rec_id[t] > nrec) nrec = rec_id[t];
    section = fmatrix(1, ntr, 1, ns);
    srcpos = fmatrix(1, 7, 1, 1);
    recpos = imatrix(1, 3, 1, nrec);
    recpos_loc = imatrix(1, 4, 1, ntr);
    for (t = 1; t <= ntr; t++) {
        for (j = 1; j <= ns; j++) section[t][j] = sec[(size_t)(t - 1) * (ns + 1) + j];
        recpos_loc[4][t] = rec_id[t - 1];
        for (j = 0; j < 3; j++) recpos[j + 1][rec_id[t - 1]] = rec_xyz[3 * (t - 1) + j];
    }
    for (j = 1; j <= 7; j++) srcpos[j][1] = 0.0f;
    if (src_xyz)
        for (j = 1; j <= 3; j++) srcpos[j][1] = src_xyz[j - 1];
    DX = dx; DY = dy; DZ = dz; DT = dt; NDT = ndt;
    outseis(stderr, fopen(path, "w"), comp, section, recpos, recpos_loc, ntr, srcpos, src_xyz ? 1 : 0, ns, 1);
    free_matrix(section, 1, ntr, 1, ns);
    free_matrix(srcpos, 1, 7, 1, 1);
    free_imatrix(recpos, 1, 3, 1, nrec);
    free_imatrix(recpos_loc, 1, 4, 1, ntr);
}

/* readseis.c for component comp of shot ishot from the file
 * "<prefix>_v?.su.shot<ishot>.0" (RUN_MULTIPLE_SHOTS = 0, MYID = 0) into
 * out [ntr][ns+1] by sample */
void rg_readseis(const char *prefix, int ishot, int comp, int ntr, int ns, float *out)
{
    int t, j;
    float **section = fmatrix(1, ntr, 1, ns), **sectionf = fmatrix(1, ntr, 1, ns);
    snprintf(SEIS_OBS_FILE, STRING_SIZE, "%s", prefix);
    RUN_MULTIPLE_SHOTS = 0;
    MYID = 0;
    if (!FP) FP = stderr;
    readseis(ishot, section, sectionf, ntr, ns, comp);
    for (t = 1; t <= ntr; t++)
        for (j = 1; j <= ns; j++) out[(size_t)(t - 1) * (ns + 1) + j] = section[t][j];
    free_matrix(section, 1, ntr, 1, ns);
    free_matrix(sectionf, 1, ntr, 1, ns);
}


/* ---- optimiser routines (lbfgs.c, lbfgs_save.c, conjugategrad.c,
 * modelupdate.c, precongrad.c, steplength.c) on compact interior arrays
 * [ny][nx][nz] (the reference's l order: j outer, i, k inner) ---------- */

static float ***cube_in(int ny, int nx, int nz, const float *a)
{
    float ***t = f3tensor(1, ny, 1, nx, 1, nz);
    int j, i, k;
    size_t l = 0;
    for (j = 1; j <= ny; j++)
        for (i = 1; i <= nx; i++)
            for (k = 1; k <= nz; k++) t[j][i][k] = a ? a[l++] : 0.0f;
    return t;
}

/* copies the cube to a (when a is not NULL) and frees it */
static void cube_out(float ***t, int ny, int nx, int nz, float *a)
{
    int j, i, k;
    size_t l = 0;
    if (a)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++)
                for (k = 1; k <= nz; k++) a[l++] = t[j][i][k];
    free_f3tensor(t, 1, ny, 1, nx, 1, nz);
}

static void opt_globals(int nx, int ny, int nz)
{
    int flag = 0;
    MPI_Initialized(&flag);
    if (!flag) MPI_Init(NULL, NULL);
    NX = nx; NY = ny; NZ = nz; NXG = nx; NYG = ny; NZG = nz;
    POS[1] = POS[2] = POS[3] = 0;
    MYID = 0;
    FP = stderr;
    FI = stderr;
}

/* lbfgs_savegrad: out [numpar * n] <- bfgsgrad1[1][1..numpar * n] */
void rg_lbfgs_savegrad(int nx, int ny, int nz, int numpar, const float *g1, const float *g2, float *out)
{
    const int n = nx * ny * nz;
    float ***a, ***b, ***c, **bg;
    int l;
    opt_globals(nx, ny, nz);
    NUMPAR = numpar;
    a = cube_in(ny, nx, nz, g1);
    b = cube_in(ny, nx, nz, g2);
    c = cube_in(ny, nx, nz, NULL);
    bg = fmatrix(1, 1, 1, numpar * n);
    lbfgs_savegrad(a, b, c, bg);
    for (l = 0; l < numpar * n; l++) out[l] = bg[1][l + 1];
    free_matrix(bg, 1, 1, 1, numpar * n);
    cube_out(a, ny, nx, nz, NULL);
    cube_out(b, ny, nx, nz, NULL);
    cube_out(c, ny, nx, nz, NULL);
}

/* lbfgs: grad1 (, grad2) in/out (the direction on return); scale
 * [bfgsnum] in/out (bfgsscale[1..bfgsnum]); mod [bfgsnum][numpar * n] in
 * (bfgsmod[1..bfgsnum]); hist [bfgsnum][numpar * n] in/out (bfgsgrad) */
void rg_lbfgs(int nx, int ny, int nz, int numpar, int bfgsnum, float *g1, float *g2, float *scale,
              const float *mod, float *hist, int iteration)
{
    const int n = nx * ny * nz, m = numpar * n;
    float ***a, ***b, ***c, **bm, **bg, *bs;
    int w, l;
    opt_globals(nx, ny, nz);
    NUMPAR = numpar;
    BFGSNUM = bfgsnum;
    a = cube_in(ny, nx, nz, g1);
    b = cube_in(ny, nx, nz, g2);
    c = cube_in(ny, nx, nz, NULL);
    bm = fmatrix(1, bfgsnum, 1, m);
    bg = fmatrix(1, bfgsnum, 1, m);
    bs = vector(1, bfgsnum);
    for (w = 1; w <= bfgsnum; w++) {
        bs[w] = scale[w - 1];
        for (l = 1; l <= m; l++) {
            bm[w][l] = mod[(size_t)(w - 1) * m + l - 1];
            bg[w][l] = hist[(size_t)(w - 1) * m + l - 1];
        }
    }
    lbfgs(a, b, c, bs, bm, bg, iteration);
    for (w = 1; w <= bfgsnum; w++) {
        scale[w - 1] = bs[w];
        for (l = 1; l <= m; l++) hist[(size_t)(w - 1) * m + l - 1] = bg[w][l];
    }
    free_vector(bs, 1, bfgsnum);
    free_matrix(bm, 1, bfgsnum, 1, m);
    free_matrix(bg, 1, bfgsnum, 1, m);
    cube_out(a, ny, nx, nz, g1);
    cube_out(b, ny, nx, nz, numpar > 1 ? g2 : NULL);
    cube_out(c, ny, nx, nz, NULL);
}

/* conjugate: g[3] in/out, prior[6] in/out, beta[3] out */
void rg_conjugate(int nx, int ny, int nz, float *g1, float *g2, float *g3, float *const prior[6], float *beta,
                  int iteration, int cdf)
{
    float ***G[3], ***P[6];
    float *gp[3] = {g1, g2, g3};
    int q;
    opt_globals(nx, ny, nz);
    for (q = 0; q < 3; q++) G[q] = cube_in(ny, nx, nz, gp[q]);
    for (q = 0; q < 6; q++) P[q] = cube_in(ny, nx, nz, prior[q]);
    conjugate(nx, ny, nz, G[0], G[1], G[2], P[0], P[1], P[2], P[3], P[4], P[5], beta, iteration, cdf);
    for (q = 0; q < 3; q++) cube_out(G[q], ny, nx, nz, gp[q]);
    for (q = 0; q < 6; q++) cube_out(P[q], ny, nx, nz, prior[q]);
}

/* modelupdate: grads (gradvp, gradvs, gradrho) in/out (normalised unless
 * lbfgs), rho/pi/u in/out; with lbfgs, mod [bfgsnum][numpar * n] in/out
 * (bfgsmod1[1..bfgsnum], slot it_group % bfgsnum written) */
void rg_modelupdate(int nx, int ny, int nz, float *g1, float *g2, float *g3, float *rho, float *pi, float *u,
                    float *mod, float step, int it_group, int lbfgs, int numpar, int bfgsnum, const float *weight,
                    float vp0, float vs0, float rho0)
{
    const int n = nx * ny * nz, m = numpar * n;
    float ***G[3], ***M[3], **bm = NULL, beta[3] = {0.0f, 0.0f, 0.0f};
    float *gp[3] = {g1, g2, g3}, *mp[3] = {rho, pi, u};
    int q, w, l;
    opt_globals(nx, ny, nz);
    LBFGS = lbfgs;
    NUMPAR = numpar;
    BFGSNUM = bfgsnum;
    WEIGHT[0] = weight[0]; WEIGHT[1] = weight[1]; WEIGHT[2] = weight[2];
    VP0 = vp0; VS0 = vs0; RHO0 = rho0;
    for (q = 0; q < 3; q++) G[q] = cube_in(ny, nx, nz, gp[q]);
    for (q = 0; q < 3; q++) M[q] = cube_in(ny, nx, nz, mp[q]);
    if (lbfgs) {
        bm = fmatrix(1, bfgsnum, 1, 2 * n);
        for (w = 1; w <= bfgsnum; w++)
            for (l = 1; l <= m; l++) bm[w][l] = mod[(size_t)(w - 1) * m + l - 1];
    }
    modelupdate(nx, ny, nz, G[0], G[1], G[2], M[0], M[1], M[2], bm, step, beta, it_group);
    if (lbfgs) {
        for (w = 1; w <= bfgsnum; w++)
            for (l = 1; l <= m; l++) mod[(size_t)(w - 1) * m + l - 1] = bm[w][l];
        free_matrix(bm, 1, bfgsnum, 1, 2 * n);
    }
    for (q = 0; q < 3; q++) cube_out(G[q], ny, nx, nz, gp[q]);
    for (q = 0; q < 3; q++) cube_out(M[q], ny, nx, nz, mp[q]);
}

/* precon_grad: grads in/out; src [nsrc][3] positions in metres (x, y, z,
 * srcpos[1..3]), rec [ntr][3] grid indices (x, y, z, recpos[1..3]) */
void rg_precon_grad(int nx, int ny, int nz, float *g1, float *g2, float *g3, int nsrc, const float *src, int ntr,
                    const int *rec, float dx, float dy, float dz, int fw, int damptype, int free_surf)
{
    float ***G[3], **srcpos = fmatrix(1, 7, 1, nsrc > 0 ? nsrc : 1);
    int **recpos = imatrix(1, 3, 1, ntr > 0 ? ntr : 1);
    float *gp[3] = {g1, g2, g3};
    int q, t;
    opt_globals(nx, ny, nz);
    DX = dx; DY = dy; DZ = dz; FW = fw; DAMPTYPE = damptype; FREE_SURF = free_surf;
    for (t = 1; t <= nsrc; t++)
        for (q = 1; q <= 3; q++) srcpos[q][t] = src[3 * (t - 1) + q - 1];
    for (t = 1; t <= ntr; t++)
        for (q = 1; q <= 3; q++) recpos[q][t] = rec[3 * (t - 1) + q - 1];
    for (q = 0; q < 3; q++) G[q] = cube_in(ny, nx, nz, gp[q]);
    precon_grad(nx, ny, nz, G[0], G[1], G[2], nsrc, srcpos, ntr, recpos, 10.0f, 1, 0);
    for (q = 0; q < 3; q++) cube_out(G[q], ny, nx, nz, gp[q]);
    free_matrix(srcpos, 1, 7, 1, nsrc > 0 ? nsrc : 1);
    free_imatrix(recpos, 1, 3, 1, ntr > 0 ? ntr : 1);
}

/* steplength: L2[4] and step[5] in/out */
void rg_steplength(float *L2, float *step, int iteration, int it_group, int lbfgs, float teststep)
{
    opt_globals(1, 1, 1);
    LBFGS = lbfgs;
    TESTSTEP = teststep;
    steplength(L2, step, iteration, it_group);
}
