// batch.hip -- the batched gradient path: several shots of one model
// propagated together on gfx950.
//
// Per time step and for all shots of the batch at once:
//   fused velocity march (interior + CPML in one pass, model shared by the
//   shots) -> point injections -> fused stress march -> explosive sources ->
//   free surface -> receivers.
// The DFT is chunked: at each sampling step the velocity march also writes
// v into a ring of `chunk` snapshots, and every `chunk` samples one pass
// reads the snapshots once and updates each accumulator `chunk` times in
// registers, so the 6*nf accumulators of a cell cross HBM once per chunk
// instead of once per sample.  The accumulator sees the same sequence of
// float roundings as one discfourier call per sample (disc_fourier.c:46-54),
// so compat results are bit-identical to the plain path.
//
// Point injections keep the reference's order around the two update
// routines: a force or adjoint source at a CPML cell is added before the
// fused update (update_v adds it before update_v_CPML updates the cell), one
// at an interior cell after it, and the adjoint traces are added a second
// time after both (update_v_ssg_CPML.c:384-393).
//
// A time loop is enqueued directly or captured once into a hipGraph and
// replayed for every batch; all per-batch inputs live in buffers the graph
// already points at.
#include <math.h>
#include <stdio.h>

#include <algorithm>
#include <atomic>
#include <map>
#include <tuple>
#include <vector>

#include "ifos_dev.h"

namespace ifos {

#define IFOS_CHUNK_MAX 32

// class of a point: 1 = in a CPML slab, 0 = in the interior box
static int point_class(const ifos_geom *g, int i, int j, int k)
{
    return (slab_h(i, g->nx, g->fw, 1) || slab_h(j, g->ny, g->fw, !g->free_surf) || slab_h(k, g->nz, g->fw, 1)) ? 1
                                                                                                              : 0;
}

struct BFields {
    float *vx, *vy, *vz, *sxx, *syy, *szz;
    int64_t fstride;
};

// forces of types 2-5 (update_v_ssg.c:474-510), one thread per shot, the
// sources of a shot in order; only sources whose cell class is `cls`
__global__ void k_bforces(DevGeom g, BFields F, const float *rip, const float *rjp, const float *rkp,
                          const int *src, const int *srccls, const float *sig, int nsrc, int nt, int ldsig, int cls,
                          Snap N)
{
    const int s = blockIdx.x;
    if (threadIdx.x != 0) return;
    const int64_t so = (int64_t)s * F.fstride;
    for (int l = 0; l < nsrc; ++l) {
        const int *p = src + 4 * ((int64_t)s * nsrc + l);
        if (srccls[(int64_t)s * nsrc + l] != cls) continue;
        const int i = p[0], j = p[1], k = p[2], type = p[3];
        if (type < 2 || type > 5) continue;
        const int64_t c = at(g, j, i, k);
        float *vx = F.vx + so, *vy = F.vy + so, *vz = F.vz + so;
        const float amp = g.dt * sig[((int64_t)s * nsrc + l) * ldsig + nt] / (g.dx * g.dy * g.dz);
        switch (type) {
        case 2: vx[c] += amp * rip[c]; break;
        case 3: vz[c] += amp * rjp[c]; break;
        case 4: vy[c] += amp * rkp[c]; break;
        case 5:
            vx[c] = (float)(vx[c] + cos(0.0) * cos(0.0) * amp);
            vy[c] = (float)(vy[c] + sin(0.0) * amp);
            vz[c] = (float)(vz[c] + sin(0.0) * cos(0.0) * amp);
            break;
        }
        if (N.ring) {
            float *r = N.ring + (int64_t)s * N.rstride + N.soff / 4;
            const int64_t fe = g.fbytes / 4;
            r[c] = vx[c];
            r[fe + c] = vy[c];
            r[2 * fe + c] = vz[c];
        }
    }
}

// adjoint traces (update_v_ssg.c:512-522): thread per (receiver, shot); the
// first receiver of a cell applies its whole same-cell chain in order.
// cls: 0 interior cells only, 1 slab cells only, 2 all.
__global__ __launch_bounds__(256) void k_badjoint(DevGeom g, BFields F, const int *rec, const int *chain,
                                                 const int *rcls, const float *adj, int ntr, int nt, int ldadj,
                                                 int cls, Snap N)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (r >= ntr) return;
    const int64_t rs = (int64_t)s * ntr;
    if (!chain[2 * (rs + r)]) return;
    if (cls != 2 && rcls[rs + r] != cls) return;
    const int *p = rec + 3 * (rs + r);
    const int64_t c = at(g, p[1], p[0], p[2]);
    const int64_t so = (int64_t)s * F.fstride;
    const int64_t comp = (int64_t)ntr * ldadj;
    const float *a = adj + (int64_t)s * 3 * comp;
    float x = F.vx[so + c], y = F.vy[so + c], z = F.vz[so + c];
    for (int q = r; q >= 0 && q < ntr; q = chain[2 * (rs + q) + 1]) {
        const int64_t o = (int64_t)q * ldadj + nt;
        x += a[o];
        z += a[2 * comp + o];
        y += a[comp + o];
    }
    F.vx[so + c] = x;
    F.vy[so + c] = y;
    F.vz[so + c] = z;
    if (N.ring) {
        float *rg = N.ring + (int64_t)s * N.rstride + N.soff / 4;
        const int64_t fe = g.fbytes / 4;
        rg[c] = x;
        rg[fe + c] = y;
        rg[2 * fe + c] = z;
    }
}

// explosive sources (psource.c:26-54), one thread per shot
__global__ void k_bpsource(DevGeom g, BFields F, const int *src, const float *sig, int nsrc, int nt, int ntall,
                           int ldsig)
{
    const int s = blockIdx.x;
    if (threadIdx.x != 0) return;
    const int64_t so = (int64_t)s * F.fstride;
    float amp = 0.0f;
    const double den = 2.0 * g.dx * g.dy * g.dz;
    for (int l = 0; l < nsrc; ++l) {
        const int *p = src + 4 * ((int64_t)s * nsrc + l);
        if (p[3] != 1) continue;
        const float *sg = sig + ((int64_t)s * nsrc + l) * ldsig;
        const int64_t c = so + at(g, p[1], p[0], p[2]);
        if (nt == 1) amp = (float)(sg[nt + 1] / den);
        if (nt > 1 && nt < ntall) amp = (float)((sg[nt + 1] - sg[nt - 1]) / den);
        if (nt == ntall) amp = (float)(-sg[nt - 1] / den);
        F.sxx[c] += amp;
        F.syy[c] += amp;
        F.szz[c] += amp;
    }
}

// seismo_ssg.c:51-53, thread per (receiver, shot); sec [S][3][ntr][ldsec]
__global__ __launch_bounds__(256) void k_bseismo(DevGeom g, BFields F, const int *rec, int ntr, int ins, int ldsec,
                                                float *sec)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (r >= ntr) return;
    const int *p = rec + 3 * ((int64_t)s * ntr + r);
    const int64_t c = (int64_t)s * F.fstride + at(g, p[1], p[0], p[2]);
    const int64_t comp = (int64_t)ntr * ldsec;
    float *o = sec + (int64_t)s * 3 * comp + (int64_t)r * ldsec + ins;
    o[0] = F.vx[c];
    o[comp] = F.vy[c];
    o[2 * comp] = F.vz[c];
}

// surface_ssg_elastic.c:78-112 for every shot (blockIdx.z)
__global__ void k_bsurface(DevGeom g, BFields F, float *sxy, float *syz, const float *u, const float *pi)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int i = blockIdx.y + 1;
    if (k > g.nz || i > g.nx) return;
    const int64_t so = (int64_t)blockIdx.z * F.fstride;
    const int64_t m = at(g, 1, i, k), c = so + m, sj = g.sj, si = g.si;
    float *sxx = F.sxx, *syy = F.syy, *szz = F.szz;
    const float *vx = F.vx, *vy = F.vy, *vz = F.vz;
    const float b1 = g.b1, b2 = g.b2;
    syy[c] = 0.0f;
    for (int q = 1; q <= 2; ++q) {
        syy[c - q * sj] = -syy[c + q * sj];
        sxy[c - q * sj] = -sxy[c + (q - 1) * sj];
        syz[c - q * sj] = -syz[c + (q - 1) * sj];
    }
    float vxx = (b1 * (vx[c] - vx[c - si]) + b2 * (vx[c + si] - vx[c - 2 * si])) / g.dx;
    float vyy = (b1 * (vy[c] - vy[c - sj]) + b2 * (vy[c + sj] - vy[c - 2 * sj])) / g.dy;
    float vzz = (b1 * (vz[c] - vz[c - 1]) + b2 * (vz[c + 1] - vz[c - 2])) / g.dz;
    float f = u[m] * 2.0f;
    float gg = pi[m];
    float h = -(g.dt * (gg - f) * (gg - f) * (vxx + vzz) / gg) - (g.dt * (gg - f) * vyy);
    sxx[c] += h;
    szz[c] += h;
}

// chunked DFT: nsnap (<= CH) snapshots of one shot's v, every accumulator
// of the cell updated nsnap times in registers.  trig[(c*NF + l)*2 + {0,1}]
// are the cos/sin factors of snapshot c (disc_fourier.c:42-44).  One thread
// per cell of [1..N]^3, shot and velocity component (blockIdx.z = shot * ny
// + j - 1, blockIdx.y = 3 * (i row block) + component), so a thread holds
// one component's snapshots and its two accumulators; the frequencies are
// walked in order with the next frequency's accumulators loaded while the
// current one is updated.
struct FBase {
    float *p[6];
};

template <int NF, int CH, bool KAHAN>
__global__ __launch_bounds__(256) void k_dft_chunk(DevGeom g, const float *ring, int64_t rstride, int nsnap,
                                                  FBase F, FBase Cc, int64_t fshot, uint32_t tot,
                                                  const double *__restrict__ trig)
{
    const int q = (int)(blockIdx.y % 3u);  // velocity component: accumulators q (Re) and 3 + q (Im)
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int i = (int)(blockIdx.y / 3u) * blockDim.y + threadIdx.y + 1;
    const int j = (int)(blockIdx.z % (unsigned)g.ny) + 1;
    const int s = (int)(blockIdx.z / (unsigned)g.ny);
    if (k > g.nz || i > g.nx) return;
    const int64_t cell = at(g, j, i, k);
    const uint32_t vo = (uint32_t)cell * 4u;
    const int64_t fe = g.fbytes / 4;
    const float *rp = ring + (int64_t)s * rstride + (int64_t)q * fe + cell;
    float v[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c)
        if (c < nsnap) v[c] = __builtin_nontemporal_load(rp + 3 * c * fe);
    rsrc_t r[2], rc[2];
#pragma unroll
    for (int n = 0; n < 2; ++n) {
        r[n] = make_rsrc(F.p[q + 3 * n] + (int64_t)s * fshot, tot);
        if (KAHAN) rc[n] = make_rsrc(Cc.p[q + 3 * n] + (int64_t)s * fshot, tot);
    }
    float a[2], cm[2], an[2], cn[2];
#pragma unroll
    for (int n = 0; n < 2; ++n) {
        a[n] = bload(r[n], vo, 0);
        cm[n] = KAHAN ? bload(rc[n], vo, 0) : 0.0f;
    }
#pragma unroll
    for (int l = 0; l < NF; ++l) {
        if (l + 1 < NF) {
#pragma unroll
            for (int n = 0; n < 2; ++n) {
                an[n] = bload(r[n], vo, (uint32_t)(l + 1) * g.fbytes);
                cn[n] = KAHAN ? bload(rc[n], vo, (uint32_t)(l + 1) * g.fbytes) : 0.0f;
            }
        }
#pragma unroll
        for (int n = 0; n < 2; ++n) {  // n = 0: cos (Re), 1: sin (Im)
            float acc = a[n], cmp = cm[n];
#pragma unroll
            for (int c = 0; c < CH; ++c)
                if (c < nsnap) {
                    const double term = (double)v[c] * trig[(c * NF + l) * 2 + n];
                    if (KAHAN) {
                        const float yv = (float)(term - (double)cmp);
                        const float t = acc + yv;
                        cmp = (t - acc) - yv;
                        acc = t;
                    } else {
                        acc = (float)(acc + term);
                    }
                }
            bstore(acc, r[n], vo, (uint32_t)l * g.fbytes);
            if (KAHAN) bstore(cmp, rc[n], vo, (uint32_t)l * g.fbytes);
        }
        if (l + 1 < NF) {
#pragma unroll
            for (int n = 0; n < 2; ++n) {
                a[n] = an[n];
                cm[n] = cn[n];
            }
        }
    }
}

static __global__ void k_bfill(float *d, int64_t n, float v)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; t < n; t += stride) d[t] = v;
}

}  // namespace ifos

using namespace ifos;

// ---------------------------------------------------------------------------
// the batch context
struct ifos_batch {
    ifos_geom g;
    int S, nf_max, ntr_max, prec, chunk;
    size_t fe;                 // elements of one field
    int64_t pe[3];             // elements of one psi_x / psi_y / psi_z array
    float *fields[9];          // [S][fe] each
    float *psi_s[9], *psi_v[9];
    // second wavefield set and velocity-derivative memory of the fused step
    // (double buffering); cur_set says which set holds the current state
    float *fields_alt[9], *psi_v_alt[9];
    int fused, fz_ti, fz_jc, cur_set, last_fused;
    int tune_fused[16][2], tune_fused_n;  // the autotuner's fused-step table
    float tune_fused_ms[16];
    float *model[9];
    float *prof;
    float *F[12], *C[12];      // [S][nf_max][fe]
    float *ring;               // [S][chunk][3][fe]
    float *sig;                // [S][nt+2]
    int *src, *srccls;         // [S][4], [S]
    int *rec, *chain, *rcls;   // [S][ntr][3], [S][ntr][2], [S][ntr]
    float *adj;                // [S][3][ntr][nt+2]
    float *syn;                // [S][3][ntr][nt+1]
    double *trig[2];           // [nsamp][nf][2] per propagation direction
    float *probe;              // [S]
    float finv[IFOS_MAX_NF];
    int nf, tast, ntast, nsamp;
    int ntr;                   // receivers of the current batch
    bool any_slab_src, any_slab_rec;
    MarchCfg cfg_v, cfg_s;     // tile shapes of the two march launches
    // viscoelastic (L mechanisms): memory variables [S][L][fe] x 6, shared
    // relaxation ratios and their averages
    int L, L_alloc;
    float eta[IFOS_MAX_L];
    float *r[6], *taus, *taup, *tavg[3];
    hipStream_t st;
    std::vector<void *> allocs;
    // captured time loops with the buffer set (cur_set) each leaves current
    std::map<std::tuple<int, int, int, int, int, int, int>, std::pair<hipGraphExec_t, int>> graphs;
    // identity of this context (never reused) and of its launch state
    // (bumped whenever cached graphs are dropped): the key of multi-level
    // graphs that include this context
    long long gen, ver;
    // the multi-level forward graph last captured with this context first
    hipGraphExec_t lv_exec;
    std::vector<long long> lv_key;
    std::vector<int> lv_sets;
};

static std::atomic<long long> g_batch_gen{0};

static int balloc(ifos_batch *b, void **p, size_t bytes)
{
    if (bytes == 0) bytes = 4;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        set_error("hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
        return IFOS_ENOMEM;
    }
    b->allocs.push_back(*p);
    return hip_check(hipMemset(*p, 0, bytes), "hipMemset");
}

static void drop_graphs(ifos_batch *b)
{
    for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second.first);
    b->graphs.clear();
    if (b->lv_exec) (void)hipGraphExecDestroy(b->lv_exec);
    b->lv_exec = nullptr;
    b->lv_key.clear();
    ++b->ver;
}

extern "C" int ifos_batch_destroy(ifos_batch *b)
{
    if (!b) return IFOS_OK;
    (void)hipStreamSynchronize(b->st);
    drop_graphs(b);
    for (void *p : b->allocs) (void)hipFree(p);
    for (int d = 0; d < 2; ++d)
        if (b->trig[d]) (void)hipFree(b->trig[d]);
    (void)hipStreamDestroy(b->st);
    delete b;
    return IFOS_OK;
}

extern "C" size_t ifos_batch_bytes(const ifos_geom *g, int nshots, int nf_max, int ntr_max, int dft_precision,
                                   int chunk)
{
    const size_t fe = ifos_field_elems(g);
    size_t psi = 0;  // (a viscoelastic batch adds 6 * L * nshots + 5 fields)
    for (int n = 0; n < 9; ++n) psi += ifos_psi_elems(g, n);
    size_t per_shot = 18 * fe + 3 * psi + 12 * (size_t)nf_max * fe * (dft_precision == IFOS_DFT_KAHAN ? 2 : 1) +
                      (size_t)chunk * 3 * fe + (size_t)(g->nt + 2) * (1 + 6 * (size_t)ntr_max);
    return 4 * ((size_t)nshots * per_shot + 9 * fe);
}

extern "C" int ifos_batch_create(ifos_batch **out, const ifos_geom *g, int nshots, int nf_max, int ntr_max,
                                 int dft_precision, int chunk)
{
    if (!out || !g || nshots < 1 || nshots > 1024 || nf_max < 1 || nf_max > 8 || ntr_max < 0 || chunk < 1 ||
        chunk > IFOS_CHUNK_MAX) {
        set_error("ifos_batch_create: bad arguments (nshots in [1,1024], nf_max in [1,8], chunk in [1,%d])",
                  IFOS_CHUNK_MAX);
        return IFOS_EINVAL;
    }
    const size_t fe = ifos_field_elems(g);
    if (3.0 * fe * 4.0 >= 4294967295.0 || (double)nf_max * fe * 4.0 >= 4294967295.0) {
        set_error("ifos_batch_create: one snapshot or the DFT array of one shot exceeds 4 GiB");
        return IFOS_EINVAL;
    }
    if ((size_t)nshots * g->ny > 65535) {
        set_error("ifos_batch_create: nshots * ny = %zu exceeds 65535", (size_t)nshots * g->ny);
        return IFOS_EINVAL;
    }
    ifos_batch *b = new ifos_batch();
    b->g = *g;
    b->S = nshots;
    b->nf_max = nf_max;
    b->ntr_max = ntr_max;
    b->prec = dft_precision;
    b->chunk = chunk;
    b->fe = fe;
    b->pe[0] = (int64_t)ifos_psi_elems(g, 0);
    b->pe[1] = (int64_t)ifos_psi_elems(g, 3);
    b->pe[2] = (int64_t)ifos_psi_elems(g, 6);
    int rc = hip_check(hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking), "hipStreamCreate");
    if (rc) {
        delete b;
        return rc;
    }
    const size_t S = nshots, nt2 = g->nt + 2, ntr = ntr_max > 0 ? ntr_max : 1;
    for (int n = 0; n < 9 && !rc; ++n) rc = balloc(b, (void **)&b->fields[n], S * fe * 4);
    for (int n = 0; n < 9 && !rc; ++n) rc = balloc(b, (void **)&b->psi_s[n], S * b->pe[n / 3] * 4);
    for (int n = 0; n < 9 && !rc; ++n) rc = balloc(b, (void **)&b->psi_v[n], S * b->pe[n / 3] * 4);
    for (int n = 0; n < 9 && !rc; ++n) rc = balloc(b, (void **)&b->fields_alt[n], S * fe * 4);
    for (int n = 0; n < 9 && !rc; ++n) rc = balloc(b, (void **)&b->psi_v_alt[n], S * b->pe[n / 3] * 4);
    for (int n = 0; n < 9 && !rc; ++n) rc = balloc(b, (void **)&b->model[n], fe * 4);
    if (!rc) rc = balloc(b, (void **)&b->prof, 18 * (size_t)(2 * g->fw + 1) * 4);
    for (int n = 0; n < 12; ++n) b->C[n] = nullptr;
    for (int n = 0; n < 12 && !rc; ++n) rc = balloc(b, (void **)&b->F[n], S * nf_max * fe * 4);
    if (dft_precision == IFOS_DFT_KAHAN)
        for (int n = 0; n < 12 && !rc; ++n) rc = balloc(b, (void **)&b->C[n], S * nf_max * fe * 4);
    if (!rc) rc = balloc(b, (void **)&b->ring, S * chunk * 3 * fe * 4);
    if (!rc) rc = balloc(b, (void **)&b->sig, S * nt2 * 4);
    if (!rc) rc = balloc(b, (void **)&b->src, S * 4 * sizeof(int));
    if (!rc) rc = balloc(b, (void **)&b->srccls, S * sizeof(int));
    if (!rc) rc = balloc(b, (void **)&b->rec, S * ntr * 3 * sizeof(int));
    if (!rc) rc = balloc(b, (void **)&b->chain, S * ntr * 2 * sizeof(int));
    if (!rc) rc = balloc(b, (void **)&b->rcls, S * ntr * sizeof(int));
    if (!rc) rc = balloc(b, (void **)&b->adj, S * 3 * ntr * nt2 * 4);
    if (!rc) rc = balloc(b, (void **)&b->syn, S * 3 * ntr * (nt2 - 1) * 4);
    if (!rc) rc = balloc(b, (void **)&b->probe, S * 4);
    b->trig[0] = b->trig[1] = nullptr;
    if (rc) {
        ifos_batch_destroy(b);
        return rc;
    }
    b->nf = 0;
    b->tast = 100;
    b->ntr = 0;
    b->cfg_v = b->cfg_s = MarchCfg{4, 32, 0, 0, 0, 1, 1, 1};  // the autotuner's usual pick
    b->fused = 0;  // opt-in; the autotuner turns it on where it beats the separate marches
    b->fz_ti = 16;
    b->fz_jc = 0;
    b->cur_set = 0;
    b->last_fused = 0;
    b->tune_fused_n = 0;
    b->gen = ++g_batch_gen;
    b->ver = 0;
    b->lv_exec = nullptr;
    b->L = 0;
    b->L_alloc = 0;
    for (int l = 0; l < IFOS_MAX_L; ++l) b->eta[l] = 0.0f;
    for (int n = 0; n < 6; ++n) b->r[n] = nullptr;
    b->taus = b->taup = nullptr;
    for (int n = 0; n < 3; ++n) b->tavg[n] = nullptr;
    *out = b;
    return IFOS_OK;
}

extern "C" int ifos_batch_set_model(ifos_batch *b, const float *d_rho, const float *d_pi, const float *d_u)
{
    const float *in[3] = {d_rho, d_pi, d_u};
    for (int n = 0; n < 3; ++n) {
        int rc = hip_check(hipMemcpyAsync(b->model[n], in[n], b->fe * 4, hipMemcpyDeviceToDevice, b->st),
                           "copy model");
        if (!rc) rc = ifos_matcopy_wrap(&b->g, b->model[n], b->st);
        if (rc) return rc;
    }
    int rc = ifos_av_mat(&b->g, b->model[0], b->model[2], b->model[3], b->model[4], b->model[5], b->model[6],
                         b->model[7], b->model[8], b->st);
    if (!rc) rc = hip_check(hipStreamSynchronize(b->st), "set_model");
    return rc;
}

// the viscoelastic update (update_s_ssg.c) with L relaxation mechanisms
// for every shot of the batch (the reference has L = 1); L = 0 switches back
// to the elastic one
extern "C" int ifos_batch_set_visco_n(ifos_batch *b, const float *d_taus, const float *d_taup, const float *eta,
                                      int L)
{
    if (!b || L < 0 || L > IFOS_MAX_L || (L >= 1 && (!d_taus || !d_taup || !eta)) ||
        (double)L * (double)b->fe * 4.0 >= 4294967296.0) {
        set_error("ifos_batch_set_visco_n: L must be 0..%d (with taus, taup and L etas; L fields under 4 GiB)",
                  IFOS_MAX_L);
        return IFOS_EINVAL;
    }
    drop_graphs(b);
    b->L = L;
    for (int l = 0; l < IFOS_MAX_L; ++l) b->eta[l] = l < L ? eta[l] : 0.0f;
    if (!L) return IFOS_OK;
    int rc = 0;
    if (!b->taus) {
        if (!rc) rc = balloc(b, (void **)&b->taus, b->fe * 4);
        if (!rc) rc = balloc(b, (void **)&b->taup, b->fe * 4);
        for (int n = 0; n < 3 && !rc; ++n) rc = balloc(b, (void **)&b->tavg[n], b->fe * 4);
        if (rc) return rc;
    }
    if (b->L_alloc < L) {  // [S][L][fe] per component (the old arrays stay with the context)
        for (int n = 0; n < 6 && !rc; ++n) rc = balloc(b, (void **)&b->r[n], (size_t)b->S * L * b->fe * 4);
        if (rc) return rc;
        b->L_alloc = L;
    }
    rc = hip_check(hipMemcpyAsync(b->taus, d_taus, b->fe * 4, hipMemcpyDeviceToDevice, b->st), "copy taus");
    if (!rc) rc = hip_check(hipMemcpyAsync(b->taup, d_taup, b->fe * 4, hipMemcpyDeviceToDevice, b->st), "copy taup");
    if (!rc) rc = ifos_matcopy_wrap(&b->g, b->taus, b->st);
    if (!rc) rc = ifos_matcopy_wrap(&b->g, b->taup, b->st);
    if (!rc) rc = ifos_av_tau(&b->g, b->taus, b->tavg[0], b->tavg[1], b->tavg[2], b->st);
    if (!rc) rc = hip_check(hipStreamSynchronize(b->st), "set_visco");
    return rc;
}

// one relaxation mechanism, as update_s_ssg.c (L = 0 or 1)
extern "C" int ifos_batch_set_visco(ifos_batch *b, const float *d_taus, const float *d_taup, float eta, int L)
{
    if (!b || L < 0 || L > 1) {
        set_error("ifos_batch_set_visco: L must be 0 or 1 (ifos_batch_set_visco_n takes more mechanisms)");
        return IFOS_EINVAL;
    }
    return ifos_batch_set_visco_n(b, d_taus, d_taup, &eta, L);
}

extern "C" int ifos_batch_set_cpml(ifos_batch *b, float vppml, float fpml)
{
    std::vector<float> h(18 * (size_t)(2 * b->g.fw + 1));
    int rc = ifos_cpml_profiles(&b->g, vppml, fpml, h.data());
    if (rc) return rc;
    return hip_check(hipMemcpy(b->prof, h.data(), h.size() * 4, hipMemcpyHostToDevice), "upload CPML profiles");
}

// trig factors of every sampling step of both directions, disc_fourier.c:36-44
extern "C" int ifos_batch_set_freqs(ifos_batch *b, const float *finv, int nf, int tast)
{
    if (nf < 1 || nf > b->nf_max) {
        set_error("ifos_batch_set_freqs: nf=%d outside [1,%d]", nf, b->nf_max);
        return IFOS_EINVAL;
    }
    const ifos_geom *g = &b->g;
    for (int l = 0; l < nf; ++l) b->finv[l] = finv[l];
    b->nf = nf;
    b->tast = tast;
    b->ntast = ifos_ntast(finv[nf - 1], tast, g->dt);
    b->nsamp = g->nt / b->ntast;
    std::vector<double> h((size_t)(b->nsamp > 0 ? b->nsamp : 1) * nf * 2);
    for (int back = 0; back < 2; ++back) {
        for (int q = 0; q < b->nsamp; ++q) {
            const int nt = (q + 1) * b->ntast;
            float t = 0.0f;
            if (back == 0) t = nt * g->dt;
            if (back == 1) t = g->time - nt * g->dt;
            for (int l = 0; l < nf; ++l) {
                h[((size_t)q * nf + l) * 2 + 0] = cos(2.0 * t * finv[l] * M_PI) * g->dt * b->ntast;
                h[((size_t)q * nf + l) * 2 + 1] = sin(2.0 * t * finv[l] * M_PI) * g->dt * b->ntast;
            }
        }
        if (b->trig[back]) (void)hipFree(b->trig[back]);
        b->trig[back] = nullptr;
        hipError_t e = hipMalloc((void **)&b->trig[back], h.size() * sizeof(double));
        if (e != hipSuccess) {
            set_error("hipMalloc(trig): %s", hipGetErrorString(e));
            return IFOS_ENOMEM;
        }
        int rc = hip_check(hipMemcpy(b->trig[back], h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice),
                           "upload trig");
        if (rc) return rc;
    }
    drop_graphs(b);  // the graphs hold the old trig pointers and sample count
    return IFOS_OK;
}

static ifos_fields bfields(const ifos_batch *b)
{
    ifos_fields f;
    float **p = (float **)&f;
    for (int n = 0; n < 9; ++n) p[n] = b->fields[n];
    return f;
}

static ifos_fields fields_of(float *const *set)
{
    ifos_fields f;
    float **p = (float **)&f;
    for (int n = 0; n < 9; ++n) p[n] = set[n];
    return f;
}

static ifos_model bmodel(const ifos_batch *b)
{
    ifos_model m;
    float **p = (float **)&m;
    for (int n = 0; n < 9; ++n) p[n] = b->model[n];
    return m;
}

static int fill(float *d, size_t n, hipStream_t st)
{
    if (n == 0) return IFOS_OK;
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_bfill, dim3(blocks), dim3(256), 0, st, d, (int64_t)n, 0.0f);
    return hip_check(hipGetLastError(), "k_bfill");
}

static int dft_chunk(ifos_batch *b, int back, int ns, int q0, int nsnap, hipStream_t st)
{
    const ifos_geom *g = &b->g;
    DevGeom G = make_dev_geom(g);
    FBase F, Cc = {};
    for (int n = 0; n < 6; ++n) {
        F.p[n] = b->F[(back ? 6 : 0) + n];
        if (b->C[0]) Cc.p[n] = b->C[(back ? 6 : 0) + n];
    }
    const int64_t rstride = (int64_t)b->chunk * 3 * b->fe;
    const int64_t fshot = (int64_t)b->nf_max * b->fe;
    const uint32_t tot = (uint32_t)((size_t)b->nf * b->fe * 4);
    const double *tr = b->trig[back] + (size_t)q0 * b->nf * 2;
    dim3 block(64, 4), grid((unsigned)((g->nz + 63) / 64), (unsigned)(3 * ((g->nx + 3) / 4)), (unsigned)(ns * g->ny));
    const bool kahan = b->prec == IFOS_DFT_KAHAN;
    ProfScope ps_(IFOS_K_DFT, st);
#define IFOS_CHUNK_LAUNCH(NFV, CHV)                                                                             \
    do {                                                                                                        \
        if (kahan)                                                                                              \
            hipLaunchKernelGGL((k_dft_chunk<NFV, CHV, true>), grid, block, 0, st, G, b->ring, rstride, nsnap, F, \
                               Cc, fshot, tot, tr);                                                             \
        else                                                                                                    \
            hipLaunchKernelGGL((k_dft_chunk<NFV, CHV, false>), grid, block, 0, st, G, b->ring, rstride, nsnap,   \
                               F, Cc, fshot, tot, tr);                                                          \
    } while (0)
#define IFOS_CHUNK_CASE(NFV)                                                                                    \
    case NFV:                                                                                                   \
        if (b->chunk <= 4) IFOS_CHUNK_LAUNCH(NFV, 4);                                                           \
        else if (b->chunk <= 8) IFOS_CHUNK_LAUNCH(NFV, 8);                                                      \
        else if (b->chunk <= 16) IFOS_CHUNK_LAUNCH(NFV, 16);                                                    \
        else if (b->chunk <= 24) IFOS_CHUNK_LAUNCH(NFV, 24);                                                    \
        else IFOS_CHUNK_LAUNCH(NFV, 32);                                                                        \
        break;
    switch (b->nf) {
        IFOS_CHUNK_CASE(1)
        IFOS_CHUNK_CASE(2)
        IFOS_CHUNK_CASE(3)
        IFOS_CHUNK_CASE(4)
        IFOS_CHUNK_CASE(5)
        IFOS_CHUNK_CASE(6)
        IFOS_CHUNK_CASE(7)
        IFOS_CHUNK_CASE(8)
    default: set_error("dft_chunk: nf=%d", b->nf); return IFOS_EINVAL;
    }
#undef IFOS_CHUNK_LAUNCH
#undef IFOS_CHUNK_CASE
    return hip_check(hipGetLastError(), "k_dft_chunk");
}

// one propagation of NT steps for shots [0, ns); back = 0 forward, 1 adjoint
static int bpass(ifos_batch *b, int back, int ns, int dft, int seis, hipStream_t st)
{
    const ifos_geom *g = &b->g;
    DevGeom G = make_dev_geom(g);
    ifos_fields f = bfields(b);
    ifos_model m = bmodel(b);
    ifos_psi ps, pv;
    for (int n = 0; n < 9; ++n) {
        ps.p[n] = b->psi_s[n];
        pv.p[n] = b->psi_v[n];
    }
    const Shots S = {ns, (int64_t)b->fe, {b->pe[0], b->pe[1], b->pe[2]}};
    const BFields BF = {f.vx, f.vy, f.vz, f.sxx, f.syy, f.szz, (int64_t)b->fe};
    const int64_t rstride = (int64_t)b->chunk * 3 * b->fe;
    const int ntr = b->ntr, nt2 = g->nt + 2;
    const dim3 rgrid((unsigned)((ntr + 255) / 256), (unsigned)ns);
    const Snap nosnap = {nullptr, 0, 0, 0};
    ViscoF V = {};
    if (b->L) {
        for (int n = 0; n < 6; ++n) V.r[n] = b->r[n];
        V.taus = b->taus;
        V.taup = b->taup;
        V.tausipjp = b->tavg[0];
        V.tausjpkp = b->tavg[1];
        V.tausipkp = b->tavg[2];
        for (int l = 0; l < IFOS_MAX_L; ++l) V.eta[l] = b->eta[l];
        V.L = b->L;
        V.lstride = (int64_t)b->fe;
    }
    int nsnap = 0, q0 = 0;
    for (int nt = 1; nt <= g->nt; ++nt) {
        const bool samp = dft && nt % b->ntast == 0 && (nt / b->ntast) <= b->nsamp;
        Snap N = nosnap;
        if (samp) N = Snap{b->ring + (size_t)nsnap * 3 * b->fe, rstride, (uint32_t)(3 * (size_t)b->fe * 4), 0};
        int rc = 0;
        // sources in CPML cells: before the fused update
        if (!back && b->any_slab_src) {
            hipLaunchKernelGGL(k_bforces, dim3(ns), dim3(64), 0, st, G, BF, m.rip, m.rjp, m.rkp, b->src, b->srccls,
                               b->sig, 1, nt, nt2, 1, nosnap);
            rc = hip_check(hipGetLastError(), "k_bforces");
        }
        if (!rc && back && b->any_slab_rec && ntr > 0) {
            hipLaunchKernelGGL(k_badjoint, rgrid, dim3(256), 0, st, G, BF, b->rec, b->chain, b->rcls, b->adj, ntr,
                               nt, nt2, 1, nosnap);
            rc = hip_check(hipGetLastError(), "k_badjoint");
        }
        if (!rc) rc = launch_v_shots(g, &f, &m, b->prof, &ps, S, samp ? &N : nullptr, b->cfg_v, st);
        if (!rc && !back) {
            ProfScope ps_(IFOS_K_SOURCES, st);
            hipLaunchKernelGGL(k_bforces, dim3(ns), dim3(64), 0, st, G, BF, m.rip, m.rjp, m.rkp, b->src, b->srccls,
                               b->sig, 1, nt, nt2, 0, N);
            rc = hip_check(hipGetLastError(), "k_bforces");
        }
        if (!rc && back && ntr > 0) {
            ProfScope ps_(IFOS_K_SOURCES, st);
            hipLaunchKernelGGL(k_badjoint, rgrid, dim3(256), 0, st, G, BF, b->rec, b->chain, b->rcls, b->adj, ntr,
                               nt, nt2, 0, N);
            hipLaunchKernelGGL(k_badjoint, rgrid, dim3(256), 0, st, G, BF, b->rec, b->chain, b->rcls, b->adj, ntr,
                               nt, nt2, 2, N);
            rc = hip_check(hipGetLastError(), "k_badjoint");
        }
        if (!rc) rc = launch_s_shots(g, &f, &m, b->prof, &pv, S, b->cfg_s, st, b->L ? &V : nullptr);
        if (!rc && !back) {
            ProfScope ps_(IFOS_K_SOURCES, st);
            hipLaunchKernelGGL(k_bpsource, dim3(ns), dim3(64), 0, st, G, BF, b->src, b->sig, 1, nt, g->nt, nt2);
            rc = hip_check(hipGetLastError(), "k_bpsource");
        }
        if (!rc && samp) {
            ++nsnap;
            if (nsnap == b->chunk) {
                rc = dft_chunk(b, back, ns, q0, nsnap, st);
                q0 += nsnap;
                nsnap = 0;
            }
        }
        if (!rc && g->free_surf) {
            hipLaunchKernelGGL(k_bsurface, dim3((unsigned)((g->nz + 63) / 64), (unsigned)g->nx, (unsigned)ns),
                               dim3(64), 0, st, G, BF, f.sxy, f.syz, m.u, m.pi);
            rc = hip_check(hipGetLastError(), "k_bsurface");
        }
        if (!rc && !back && seis && ntr > 0) {
            ProfScope ps_(IFOS_K_SOURCES, st);
            hipLaunchKernelGGL(k_bseismo, rgrid, dim3(256), 0, st, G, BF, b->rec, ntr, nt, g->nt + 1, b->syn);
            rc = hip_check(hipGetLastError(), "k_bseismo");
        }
        if (rc) return rc;
    }
    if (nsnap > 0) return dft_chunk(b, back, ns, q0, nsnap, st);
    return IFOS_OK;
}

// the fused form of bpass: v(1) alone, then one fused launch per step doing
// sigma(n) + psource(n) + v(n+1) from one buffer set into the other.  The
// stress of the last step feeds nothing the pass returns and is not formed.
// Needs no free surface, no point in a CPML slab and the elastic update.
static bool fused_ok(const ifos_batch *b)
{
    return b->fused && !b->g.free_surf && !b->any_slab_src && !b->any_slab_rec && !b->L;
}

static int bpass_fused(ifos_batch *b, int back, int ns, int dft, int seis, hipStream_t st)
{
    const ifos_geom *g = &b->g;
    DevGeom G = make_dev_geom(g);
    ifos_model m = bmodel(b);
    float *cur[9], *nxt[9], *pvc[9], *pvn[9];
    for (int n = 0; n < 9; ++n) {
        cur[n] = b->fields[n];
        nxt[n] = b->fields_alt[n];
        pvc[n] = b->psi_v[n];
        pvn[n] = b->psi_v_alt[n];
    }
    ifos_psi ps;
    for (int n = 0; n < 9; ++n) ps.p[n] = b->psi_s[n];
    const Shots S = {ns, (int64_t)b->fe, {b->pe[0], b->pe[1], b->pe[2]}};
    const int64_t rstride = (int64_t)b->chunk * 3 * b->fe;
    const int ntr = b->ntr, nt2 = g->nt + 2;
    const dim3 rgrid((unsigned)((ntr + 255) / 256), (unsigned)ns);
    int nsnap = 0, q0 = 0, set = 0;
    for (int nt = 1; nt <= g->nt; ++nt) {
        const bool samp = dft && nt % b->ntast == 0 && (nt / b->ntast) <= b->nsamp;
        Snap N = {nullptr, 0, 0, 0};
        if (samp) N = Snap{b->ring + (size_t)nsnap * 3 * b->fe, rstride, (uint32_t)(3 * (size_t)b->fe * 4), 0};
        int rc = 0;
        if (nt == 1) {
            ifos_fields f = fields_of(cur);
            rc = launch_v_shots(g, &f, &m, b->prof, &ps, S, samp ? &N : nullptr, b->cfg_v, st);
        } else {
            // sigma(nt-1) + psource(nt-1) + v(nt): cur -> nxt
            FusedArgs A;
            A.vx = cur[0]; A.vy = cur[1]; A.vz = cur[2];
            A.sxx = cur[3]; A.syy = cur[4]; A.szz = cur[5]; A.sxy = cur[6]; A.syz = cur[7]; A.sxz = cur[8];
            A.nvx = nxt[0]; A.nvy = nxt[1]; A.nvz = nxt[2];
            A.nsxx = nxt[3]; A.nsyy = nxt[4]; A.nszz = nxt[5]; A.nsxy = nxt[6]; A.nsyz = nxt[7]; A.nsxz = nxt[8];
            A.pi = m.pi; A.u = m.u; A.uipjp = m.uipjp; A.ujpkp = m.ujpkp; A.uipkp = m.uipkp;
            A.rip = m.rip; A.rjp = m.rjp; A.rkp = m.rkp;
            for (int n = 0; n < 9; ++n) {
                A.psv_old.p[n] = pvc[n];
                A.psv_new.p[n] = pvn[n];
                A.pss.p[n] = b->psi_s[n];
            }
            A.src = back ? nullptr : b->src;
            A.sig = b->sig;
            A.ldsig = nt2;
            A.nt = nt - 1;
            A.ntall = g->nt;
            rc = launch_vs_fused(g, A, b->prof, S, samp ? &N : nullptr, b->fz_ti, b->fz_jc, st);
            for (int n = 0; n < 9; ++n) {
                std::swap(cur[n], nxt[n]);
                std::swap(pvc[n], pvn[n]);
            }
            set ^= 1;
        }
        const BFields BF = {cur[0], cur[1], cur[2], cur[3], cur[4], cur[5], (int64_t)b->fe};
        if (!rc && !back) {
            ProfScope ps_(IFOS_K_SOURCES, st);
            hipLaunchKernelGGL(k_bforces, dim3(ns), dim3(64), 0, st, G, BF, m.rip, m.rjp, m.rkp, b->src, b->srccls,
                               b->sig, 1, nt, nt2, 0, N);
            rc = hip_check(hipGetLastError(), "k_bforces");
        }
        if (!rc && back && ntr > 0) {
            ProfScope ps_(IFOS_K_SOURCES, st);
            hipLaunchKernelGGL(k_badjoint, rgrid, dim3(256), 0, st, G, BF, b->rec, b->chain, b->rcls, b->adj, ntr,
                               nt, nt2, 0, N);
            hipLaunchKernelGGL(k_badjoint, rgrid, dim3(256), 0, st, G, BF, b->rec, b->chain, b->rcls, b->adj, ntr,
                               nt, nt2, 2, N);
            rc = hip_check(hipGetLastError(), "k_badjoint");
        }
        if (!rc && samp) {
            ++nsnap;
            if (nsnap == b->chunk) {
                rc = dft_chunk(b, back, ns, q0, nsnap, st);
                q0 += nsnap;
                nsnap = 0;
            }
        }
        if (!rc && !back && seis && ntr > 0) {
            ProfScope ps_(IFOS_K_SOURCES, st);
            hipLaunchKernelGGL(k_bseismo, rgrid, dim3(256), 0, st, G, BF, b->rec, ntr, nt, g->nt + 1, b->syn);
            rc = hip_check(hipGetLastError(), "k_bseismo");
        }
        if (rc) return rc;
    }
    b->cur_set = set;
    if (nsnap > 0) return dft_chunk(b, back, ns, q0, nsnap, st);
    return IFOS_OK;
}

static int zero_state(ifos_batch *b, int ns, hipStream_t st)
{
    int rc = 0;
    for (int n = 0; n < 9 && !rc; ++n) rc = fill(b->fields[n], (size_t)ns * b->fe, st);
    for (int n = 0; n < 9 && !rc; ++n) rc = fill(b->psi_s[n], (size_t)ns * b->pe[n / 3], st);
    for (int n = 0; n < 9 && !rc; ++n) rc = fill(b->psi_v[n], (size_t)ns * b->pe[n / 3], st);
    for (int n = 0; n < 9 && !rc; ++n) rc = fill(b->fields_alt[n], (size_t)ns * b->fe, st);
    for (int n = 0; n < 9 && !rc; ++n) rc = fill(b->psi_v_alt[n], (size_t)ns * b->pe[n / 3], st);
    b->cur_set = 0;
    if (b->L)  // zero_wavefield.c:80-95
        for (int n = 0; n < 6 && !rc; ++n) rc = fill(b->r[n], (size_t)ns * b->L * b->fe, st);
    return rc;
}

// runs the pass, through a cached graph when use_graph
static int run_pass(ifos_batch *b, int back, int ns, int dft, int seis, int use_graph)
{
    hipStream_t st = b->st;
    const bool fz = fused_ok(b);
    b->last_fused = fz;
    auto pass = [&](hipStream_t s) { return fz ? bpass_fused(b, back, ns, dft, seis, s) : bpass(b, back, ns, dft, seis, s); };
    if (!use_graph) return pass(st);
    auto key = std::make_tuple(back, ns, dft, seis, b->ntr, (int)b->any_slab_src, (int)b->any_slab_rec);
    auto it = b->graphs.find(key);
    if (it == b->graphs.end()) {
        hipGraph_t graph;
        int rc = hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        if (rc) return rc;
        int rc2 = pass(st);
        rc = hip_check(hipStreamEndCapture(st, &graph), "hipStreamEndCapture");
        if (rc2) return rc2;
        if (rc) return rc;
        hipGraphExec_t exec;
        rc = hip_check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
        (void)hipGraphDestroy(graph);
        if (rc) return rc;
        it = b->graphs.emplace(key, std::make_pair(exec, b->cur_set)).first;
    }
    b->cur_set = it->second.second;  // the replayed loop ends in the set the captured one did
    return hip_check(hipGraphLaunch(it->second.first, st), "hipGraphLaunch");
}

static int check_probe(ifos_batch *b, int ns, const char *what)
{
    const ifos_geom *g = &b->g;
    const size_t off = ifos_offset(g, g->ny / 2, g->nx / 2, g->nz / 2);
    std::vector<float> h(ns);
    for (int s = 0; s < ns; ++s) {
        const float *vy = b->cur_set ? b->fields_alt[1] : b->fields[1];
        int rc = hip_check(hipMemcpyAsync(&h[s], vy + (size_t)s * b->fe + off, 4, hipMemcpyDeviceToHost,
                                          b->st),
                           "probe");
        if (rc) return rc;
    }
    int rc = hip_check(hipStreamSynchronize(b->st), "probe sync");
    if (rc) return rc;
    for (int s = 0; s < ns; ++s)
        if (isnan(h[s])) {
            set_error("simulation unstable: NaN at the probe cell after the %s pass (shot %d of the batch)", what, s);
            return IFOS_ENAN;
        }
    return IFOS_OK;
}

// checks the shots of one batch, uploads their sources and receivers and
// zeroes the state they start from (ifos3d.c:727-739)
static int batch_prepare(ifos_batch *b, const ifos_shot *shots, int ns, const char *who)
{
    if (!b || !shots || ns < 1 || ns > b->S) {
        set_error("%s: need 1..%d shots", who, b ? b->S : 0);
        return IFOS_EINVAL;
    }
    const ifos_geom *g = &b->g;
    const int want_adj = shots[0].want_adjoint;
    const int ntr = shots[0].ntr;
    for (int s = 0; s < ns; ++s) {
        const ifos_shot &sh = shots[s];
        if (sh.want_adjoint != want_adj || sh.ntr != ntr) {
            set_error("%s: shots of one batch need the same ntr and want_adjoint", who);
            return IFOS_EINVAL;
        }
        if (sh.ntr > b->ntr_max) {
            set_error("%s: ntr=%d > ntr_max=%d", who, sh.ntr, b->ntr_max);
            return IFOS_EINVAL;
        }
        if (sh.src_i < 1 || sh.src_i > g->nx || sh.src_j < 1 || sh.src_j > g->ny || sh.src_k < 1 || sh.src_k > g->nz) {
            set_error("%s: source of shot %d outside the grid", who, s);
            return IFOS_EINVAL;
        }
        for (int r = 0; r < sh.ntr; ++r) {
            const int *p = sh.rec_ijk + 3 * r;
            if (p[0] < 1 || p[0] > g->nx || p[1] < 1 || p[1] > g->ny || p[2] < 1 || p[2] > g->nz) {
                set_error("%s: receiver %d of shot %d outside the grid", who, r, s);
                return IFOS_EINVAL;
            }
        }
        if (want_adj && !sh.d_obs) {
            set_error("%s: shot %d has no observed data", who, s);
            return IFOS_EINVAL;
        }
    }
    if (want_adj && b->nf < 1) {
        set_error("%s: frequencies not set", who);
        return IFOS_EINVAL;
    }
    hipStream_t st = b->st;
    const size_t nt2 = g->nt + 2, ntrs = ntr > 0 ? ntr : 1;
    // per-shot inputs (host -> device, once per batch)
    std::vector<float> sig((size_t)ns * nt2);
    std::vector<int> src(4 * (size_t)ns), srccls(ns), rec(3 * (size_t)ns * ntrs), chain(2 * (size_t)ns * ntrs),
        rcls((size_t)ns * ntrs);
    b->any_slab_src = b->any_slab_rec = false;
    for (int s = 0; s < ns; ++s) {
        const ifos_shot &sh = shots[s];
        int rc = ifos_wavelet(g, sh.source_shape, sh.tshift, sh.fc, sh.amp, sig.data() + (size_t)s * nt2);
        if (rc) return rc;
        src[4 * s] = sh.src_i;
        src[4 * s + 1] = sh.src_j;
        src[4 * s + 2] = sh.src_k;
        src[4 * s + 3] = sh.src_type;
        srccls[s] = point_class(g, sh.src_i, sh.src_j, sh.src_k);
        if (srccls[s] && sh.src_type >= 2 && sh.src_type <= 5) b->any_slab_src = true;
        if (ntr > 0) {
            rc = ifos_rec_chain(sh.rec_ijk, ntr, chain.data() + 2 * (size_t)s * ntr);
            if (rc) return rc;
            for (int r = 0; r < ntr; ++r) {
                const int *p = sh.rec_ijk + 3 * r;
                for (int d = 0; d < 3; ++d) rec[3 * ((size_t)s * ntr + r) + d] = p[d];
                rcls[(size_t)s * ntr + r] = point_class(g, p[0], p[1], p[2]);
                if (rcls[(size_t)s * ntr + r]) b->any_slab_rec = true;
            }
        }
    }
    b->ntr = ntr;
    int rc = hip_check(hipMemcpyAsync(b->sig, sig.data(), sig.size() * 4, hipMemcpyHostToDevice, st), "upload sig");
    if (!rc) rc = hip_check(hipMemcpyAsync(b->src, src.data(), src.size() * 4, hipMemcpyHostToDevice, st), "upload src");
    if (!rc)
        rc = hip_check(hipMemcpyAsync(b->srccls, srccls.data(), srccls.size() * 4, hipMemcpyHostToDevice, st),
                       "upload srccls");
    if (!rc && ntr > 0) {
        rc = hip_check(hipMemcpyAsync(b->rec, rec.data(), rec.size() * 4, hipMemcpyHostToDevice, st), "upload rec");
        if (!rc)
            rc = hip_check(hipMemcpyAsync(b->chain, chain.data(), chain.size() * 4, hipMemcpyHostToDevice, st),
                           "upload chain");
        if (!rc)
            rc = hip_check(hipMemcpyAsync(b->rcls, rcls.data(), rcls.size() * 4, hipMemcpyHostToDevice, st),
                           "upload rcls");
    }
    // ifos3d.c:738-739: zero wavefields and the accumulators of these shots
    if (!rc) rc = zero_state(b, ns, st);
    const size_t fshot = (size_t)b->nf_max * b->fe;
    for (int n = 0; n < 12 && !rc && want_adj; ++n) {
        rc = fill(b->F[n], (size_t)ns * fshot, st);
        if (!rc && b->C[n]) rc = fill(b->C[n], (size_t)ns * fshot, st);
    }
    const size_t synshot = 3 * (size_t)ntr * (g->nt + 1);
    if (!rc) rc = fill(b->syn, (size_t)ns * synshot, st);
    if (!rc) rc = hip_check(hipStreamSynchronize(st), "batch uploads");
    return rc;
}

extern "C" int ifos_batch_run(ifos_batch *b, const ifos_shot *shots, int ns, float *d_syn,
                              float *const d_grad[3], double *l2, int use_graph)
{
    int rc = batch_prepare(b, shots, ns, "ifos_batch_run");
    if (rc) return rc;
    const ifos_geom *g = &b->g;
    hipStream_t st = b->st;
    const int want_adj = shots[0].want_adjoint;
    const int ntr = shots[0].ntr;
    const size_t nt2 = g->nt + 2;
    const size_t fshot = (size_t)b->nf_max * b->fe;
    const size_t synshot = 3 * (size_t)ntr * (g->nt + 1);

    rc = run_pass(b, 0, ns, want_adj, 1, use_graph);
    if (!rc) rc = check_probe(b, ns, "forward");
    if (!rc && d_syn)
        rc = hip_check(hipMemcpyAsync(d_syn, b->syn, (size_t)ns * synshot * 4, hipMemcpyDeviceToDevice, st),
                       "copy seismograms");
    if (rc || !want_adj) {
        if (!rc) rc = hip_check(hipStreamSynchronize(st), "batch forward");
        return rc;
    }
    // residual -> adjoint sources (ifos3d.c:837-866)
    const size_t adjshot = 3 * (size_t)ntr * nt2;
    rc = fill(b->adj, (size_t)ns * adjshot, st);
    for (int s = 0; s < ns && !rc; ++s) {
        double l2s = 0.0;
        rc = ifos_residual(g, shots[s].d_obs, b->syn + s * synshot, b->adj + s * adjshot, ntr, g->nt, &l2s, st);
        if (l2) l2[s] = l2s;
    }
    if (!rc) rc = zero_state(b, ns, st);
    if (!rc) rc = run_pass(b, 1, ns, 1, 0, use_graph);
    if (!rc) rc = check_probe(b, ns, "adjoint");
    if (rc) return rc;
    ifos_model m = bmodel(b);
    float *Ff[6], *Fb[6];
    for (int n = 0; n < 6; ++n) {
        Ff[n] = b->F[n];
        Fb[n] = b->F[6 + n];
    }
    if (d_grad) rc = launch_gradient_shots(g, Ff, Fb, (int64_t)fshot, ns, &m, b->finv, b->nf, d_grad, st);
    if (!rc) rc = hip_check(hipStreamSynchronize(st), "batch gradient");
    return rc;
}

// ---------------------------------------------------------------------------
// grid levels of uneven size run together: each level is a batch context of
// its own grid, and the forward time loops of all levels are enqueued on the
// levels' own streams, or captured once into one graph whose branches are
// those loops (fork from the first level's stream, join back into it) and
// replayed.  A level's shots give the same seismograms as ifos_batch_run on
// that level alone.
static int level_pass(ifos_batch *b, int ns, hipStream_t s)
{
    const bool fz = fused_ok(b);
    b->last_fused = fz;
    return fz ? bpass_fused(b, 0, ns, 0, 1, s) : bpass(b, 0, ns, 0, 1, s);
}

// captures the forward loops of all levels into one graph
static int capture_levels(ifos_batch *const *lv, int nlev, const int *ns, hipGraphExec_t *exec)
{
    hipStream_t s0 = lv[0]->st;
    std::vector<hipEvent_t> ev(nlev, nullptr);
    int rc = IFOS_OK;
    for (int l = 0; l < nlev && !rc; ++l)
        rc = hip_check(hipEventCreateWithFlags(&ev[l], hipEventDisableTiming), "hipEventCreate");
    if (rc) {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        return rc;
    }
    rc = hip_check(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    if (rc) {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        return rc;
    }
    // fork: every other level's stream joins the capture behind s0
    int rc2 = hip_check(hipEventRecord(ev[0], s0), "fork");
    std::vector<bool> joined(nlev, false);
    for (int l = 1; l < nlev && !rc2; ++l) {
        rc2 = hip_check(hipStreamWaitEvent(lv[l]->st, ev[0], 0), "fork wait");
        joined[l] = !rc2;
    }
    for (int l = 0; l < nlev && !rc2; ++l) rc2 = level_pass(lv[l], ns[l], lv[l]->st);
    // join every stream that entered the capture, also after an error, so
    // that the capture can end
    for (int l = 1; l < nlev; ++l)
        if (joined[l]) {
            int r = hip_check(hipEventRecord(ev[l], lv[l]->st), "join");
            if (!r) r = hip_check(hipStreamWaitEvent(s0, ev[l], 0), "join wait");
            if (!rc2) rc2 = r;
        }
    hipGraph_t graph = nullptr;
    rc = hip_check(hipStreamEndCapture(s0, &graph), "hipStreamEndCapture");
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    if (!rc && !rc2) rc = hip_check(hipGraphInstantiate(exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    if (graph) (void)hipGraphDestroy(graph);
    return rc2 ? rc2 : rc;
}

extern "C" int ifos_levels_run(ifos_batch *const *lv, int nlev, const ifos_shot *const *shots, const int *ns,
                               float *const *d_syn, int use_graph)
{
    if (!lv || !shots || !ns || nlev < 1 || nlev > IFOS_MAX_LEVELS) {
        set_error("ifos_levels_run: need 1..%d levels", IFOS_MAX_LEVELS);
        return IFOS_EINVAL;
    }
    for (int l = 0; l < nlev; ++l) {
        if (!lv[l] || !shots[l]) {
            set_error("ifos_levels_run: level %d has no context or no shots", l);
            return IFOS_EINVAL;
        }
        for (int m = 0; m < l; ++m)
            if (lv[m] == lv[l]) {
                set_error("ifos_levels_run: levels %d and %d share one context", m, l);
                return IFOS_EINVAL;
            }
        for (int s = 0; s < ns[l] && s < lv[l]->S; ++s)
            if (shots[l][s].want_adjoint) {
                set_error("ifos_levels_run: forward solves only (shot %d of level %d asks for the adjoint)", s, l);
                return IFOS_EINVAL;
            }
    }
    int rc = IFOS_OK;
    for (int l = 0; l < nlev && !rc; ++l) rc = batch_prepare(lv[l], shots[l], ns[l], "ifos_levels_run");
    if (rc) return rc;
    if (!use_graph) {
        for (int l = 0; l < nlev && !rc; ++l) rc = level_pass(lv[l], ns[l], lv[l]->st);
        for (int l = 0; l < nlev; ++l) {
            int r = hip_check(hipStreamSynchronize(lv[l]->st), "levels forward");
            if (!rc) rc = r;
        }
    } else {
        std::vector<long long> key;
        for (int l = 0; l < nlev; ++l) {
            const ifos_batch *b = lv[l];
            key.insert(key.end(), {b->gen, b->ver, ns[l], b->ntr, (long long)b->any_slab_src,
                                   (long long)b->any_slab_rec, (long long)fused_ok(b)});
        }
        ifos_batch *b0 = lv[0];
        if (!b0->lv_exec || b0->lv_key != key) {
            if (b0->lv_exec) (void)hipGraphExecDestroy(b0->lv_exec);
            b0->lv_exec = nullptr;
            b0->lv_key.clear();
            hipGraphExec_t exec = nullptr;
            rc = capture_levels(lv, nlev, ns, &exec);
            if (rc) return rc;
            b0->lv_exec = exec;
            b0->lv_key = key;
            b0->lv_sets.resize(nlev);
            for (int l = 0; l < nlev; ++l) b0->lv_sets[l] = lv[l]->cur_set;
        }
        for (int l = 0; l < nlev; ++l) lv[l]->cur_set = b0->lv_sets[l];
        rc = hip_check(hipGraphLaunch(b0->lv_exec, b0->st), "hipGraphLaunch");
        if (!rc) rc = hip_check(hipStreamSynchronize(b0->st), "levels forward");
    }
    for (int l = 0; l < nlev && !rc; ++l) {
        ifos_batch *b = lv[l];
        rc = check_probe(b, ns[l], "forward");
        if (!rc && d_syn && d_syn[l]) {
            const size_t synshot = 3 * (size_t)b->ntr * (b->g.nt + 1);
            rc = hip_check(hipMemcpyAsync(d_syn[l], b->syn, (size_t)ns[l] * synshot * 4, hipMemcpyDeviceToDevice,
                                          b->st),
                           "copy seismograms");
        }
        if (!rc) rc = hip_check(hipStreamSynchronize(b->st), "levels seismograms");
    }
    return rc;
}

// ---------------------------------------------------------------------------
// autotuner: times each candidate tile shape of the two march launches on
// the batch's own buffers (ns shots, `reps` launches each) and keeps the
// fastest.  The wavefields are zeroed afterwards.
static const MarchCfg k_candidates[] = {
    {4, 32, 0, 0, 0, 1, 1, 0}, {4, 32, 0, 0, 1, 1, 1, 0}, {4, 0, 0, 0, 0, 1, 1, 0}, {8, 32, 0, 0, 0, 1, 1, 0},
    {2, 0, 0, 0, 1, 1, 1, 0},  {1, 0, 0, 0, 1, 1, 1, 0},  {4, 0, 0, 1, 1, 1, 1, 0}, {4, 32, 0, 0, 0, 1, 2, 0},
    {4, 32, 0, 0, 1, 1, 2, 0}, {8, 32, 0, 0, 0, 1, 2, 0}, {4, 32, 0, 0, 0, 1, 4, 0}, {2, 32, 0, 0, 1, 1, 2, 0},
    {4, 32, 0, 0, 0, 1, 1, 1}, {8, 32, 0, 0, 0, 1, 1, 1}, {2, 32, 0, 0, 0, 1, 1, 1}, {4, 0, 0, 0, 0, 1, 1, 1},
    {8, 0, 0, 0, 0, 1, 1, 1},  {4, 64, 0, 0, 0, 1, 1, 1}, {8, 16, 0, 0, 0, 1, 1, 1}, {2, 0, 0, 0, 0, 1, 1, 1},
};

extern "C" int ifos_batch_autotune(ifos_batch *b, int ns, int reps, int *best_v, int *best_s, float *ms_v,
                                   float *ms_s)
{
    if (!b || ns < 1 || ns > b->S || reps < 1) return IFOS_EINVAL;
    const int nc = (int)(sizeof(k_candidates) / sizeof(k_candidates[0]));
    hipStream_t st = b->st;
    const ifos_geom *g = &b->g;
    ifos_fields f = bfields(b);
    ifos_model m = bmodel(b);
    ifos_psi ps, pv;
    for (int n = 0; n < 9; ++n) {
        ps.p[n] = b->psi_s[n];
        pv.p[n] = b->psi_v[n];
    }
    const Shots S = {ns, (int64_t)b->fe, {b->pe[0], b->pe[1], b->pe[2]}};
    ViscoF V = {};
    if (b->L) {
        for (int n = 0; n < 6; ++n) V.r[n] = b->r[n];
        V.taus = b->taus;
        V.taup = b->taup;
        V.tausipjp = b->tavg[0];
        V.tausjpkp = b->tavg[1];
        V.tausipkp = b->tavg[2];
        for (int l = 0; l < IFOS_MAX_L; ++l) V.eta[l] = b->eta[l];
        V.L = b->L;
        V.lstride = (int64_t)b->fe;
    }
    hipEvent_t e0, e1;
    int rc = hip_check(hipEventCreate(&e0), "event");
    if (!rc) rc = hip_check(hipEventCreate(&e1), "event");
    if (rc) return rc;
    float bv = 1e30f, bs = 1e30f;
    int iv = 0, is = 0;
    for (int c = 0; c < nc && !rc; ++c) {
        for (int kind = 0; kind < 2 && !rc; ++kind) {
            // one untimed launch, then reps timed
            for (int r = 0; r <= reps && !rc; ++r) {
                if (r == 1) rc = hip_check(hipEventRecord(e0, st), "record");
                if (!rc)
                    rc = kind == 0 ? launch_v_shots(g, &f, &m, b->prof, &ps, S, nullptr, k_candidates[c], st)
                                   : launch_s_shots(g, &f, &m, b->prof, &pv, S, k_candidates[c], st,
                                                    b->L ? &V : nullptr);
            }
            if (!rc) rc = hip_check(hipEventRecord(e1, st), "record");
            if (!rc) rc = hip_check(hipEventSynchronize(e1), "sync");
            float ms = 0.0f;
            if (!rc) rc = hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
            ms /= reps;
            if (ms_v && kind == 0) ms_v[c] = ms;
            if (ms_s && kind == 1) ms_s[c] = ms;
            if (kind == 0 && ms < bv) { bv = ms; iv = c; }
            if (kind == 1 && ms < bs) { bs = ms; is = c; }
        }
    }
    // the fused step: (ti, jc) candidates, timed like the others
    static const int fz[][2] = {{16, 0}, {16, 16}, {16, 32}, {16, 64}, {8, 0}, {8, 16}, {8, 32}};
    float bf = 1e30f;
    int ifz = 0;
    b->tune_fused_n = 0;
    if (!rc && !b->g.free_surf && !b->L) {
        FusedArgs A;
        A.vx = b->fields[0]; A.vy = b->fields[1]; A.vz = b->fields[2];
        A.sxx = b->fields[3]; A.syy = b->fields[4]; A.szz = b->fields[5];
        A.sxy = b->fields[6]; A.syz = b->fields[7]; A.sxz = b->fields[8];
        A.nvx = b->fields_alt[0]; A.nvy = b->fields_alt[1]; A.nvz = b->fields_alt[2];
        A.nsxx = b->fields_alt[3]; A.nsyy = b->fields_alt[4]; A.nszz = b->fields_alt[5];
        A.nsxy = b->fields_alt[6]; A.nsyz = b->fields_alt[7]; A.nsxz = b->fields_alt[8];
        A.pi = m.pi; A.u = m.u; A.uipjp = m.uipjp; A.ujpkp = m.ujpkp; A.uipkp = m.uipkp;
        A.rip = m.rip; A.rjp = m.rjp; A.rkp = m.rkp;
        for (int n = 0; n < 9; ++n) {
            A.psv_old.p[n] = b->psi_v[n];
            A.psv_new.p[n] = b->psi_v_alt[n];
            A.pss.p[n] = b->psi_s[n];
        }
        A.src = nullptr;
        A.sig = b->sig;
        A.ldsig = g->nt + 2;
        A.nt = 1;
        A.ntall = g->nt;
        const int nfz = (int)(sizeof(fz) / sizeof(fz[0]));
        for (int c = 0; c < nfz && !rc; ++c) {
            for (int r = 0; r <= reps && !rc; ++r) {
                if (r == 1) rc = hip_check(hipEventRecord(e0, st), "record");
                if (!rc) rc = launch_vs_fused(g, A, b->prof, S, nullptr, fz[c][0], fz[c][1], st);
            }
            if (!rc) rc = hip_check(hipEventRecord(e1, st), "record");
            if (!rc) rc = hip_check(hipEventSynchronize(e1), "sync");
            float ms = 0.0f;
            if (!rc) rc = hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
            ms /= reps;
            b->tune_fused[c][0] = fz[c][0];
            b->tune_fused[c][1] = fz[c][1];
            b->tune_fused_ms[c] = ms;
            b->tune_fused_n = c + 1;
            if (ms < bf) { bf = ms; ifz = c; }
        }
        if (!rc) {
            b->fz_ti = fz[ifz][0];
            b->fz_jc = fz[ifz][1];
            // the fused step replaces one velocity and one stress launch
            b->fused = bf < bv + bs;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    b->cfg_v = k_candidates[iv];
    b->cfg_s = k_candidates[is];
    drop_graphs(b);
    if (best_v) *best_v = iv;
    if (best_s) *best_s = is;
    rc = zero_state(b, b->S, st);
    if (!rc) rc = hip_check(hipStreamSynchronize(st), "autotune");
    return rc;
}

extern "C" int ifos_batch_tune_candidates(int *cfg, int max)
{
    const int nc = (int)(sizeof(k_candidates) / sizeof(k_candidates[0]));
    for (int c = 0; c < nc && c < max; ++c) {
        int *o = cfg + IFOS_TILE_FIELDS * c;
        o[0] = k_candidates[c].ti;
        o[1] = k_candidates[c].jc;
        o[2] = k_candidates[c].group;
        o[3] = k_candidates[c].split;
        o[4] = k_candidates[c].pipe;
        o[5] = k_candidates[c].kb;
        o[6] = k_candidates[c].sb;
        o[7] = k_candidates[c].lds;
    }
    return nc;
}

extern "C" int ifos_batch_set_tiles(ifos_batch *b, const int cfg_v[IFOS_TILE_FIELDS],
                                    const int cfg_s[IFOS_TILE_FIELDS])
{
    if (!b) return IFOS_EINVAL;
    for (const int *c : {cfg_v, cfg_s})
        if (c && check_tile_cfg(c, "ifos_batch_set_tiles")) return IFOS_EINVAL;
    if (cfg_v) b->cfg_v = MarchCfg{cfg_v[0], cfg_v[1], cfg_v[2], cfg_v[3], cfg_v[4], cfg_v[5], cfg_v[6], cfg_v[7]};
    if (cfg_s) b->cfg_s = MarchCfg{cfg_s[0], cfg_s[1], cfg_s[2], cfg_s[3], cfg_s[4], cfg_s[5], cfg_s[6], cfg_s[7]};
    drop_graphs(b);
    return IFOS_OK;
}

// the fused step on (when the configuration allows it) or off, its
// workgroup rows (8 or 16) and j chunk (<= 0: automatic)
extern "C" int ifos_batch_set_fused(ifos_batch *b, int on, int ti, int jc)
{
    if (!b || (ti != 8 && ti != 16)) {
        set_error("ifos_batch_set_fused: ti must be 8 or 16");
        return IFOS_EINVAL;
    }
    b->fused = on != 0;
    b->fz_ti = ti;
    b->fz_jc = jc;
    drop_graphs(b);
    return IFOS_OK;
}

// the autotuner's fused-step timings: out[3*c] = ti, jc, ms*1000; returns count
extern "C" int ifos_batch_tune_fused(const ifos_batch *b, float *out, int max)
{
    if (!b) return 0;
    int n = 0;
    for (; n < b->tune_fused_n && n < max; ++n) {
        out[3 * n] = (float)b->tune_fused[n][0];
        out[3 * n + 1] = (float)b->tune_fused[n][1];
        out[3 * n + 2] = b->tune_fused_ms[n];
    }
    return n;
}

// 1 if the last run used the fused step
extern "C" int ifos_batch_last_fused(const ifos_batch *b) { return b ? b->last_fused : 0; }

// DFT accumulators of one shot of the last batch (device pointers into the
// batch's arrays; frequency l of component n at d_F[n] + l * field elements)
extern "C" int ifos_batch_dft(ifos_batch *b, int shot, int back, float *d_F[6])
{
    if (!b || shot < 0 || shot >= b->S) return IFOS_EINVAL;
    for (int n = 0; n < 6; ++n) d_F[n] = b->F[(back ? 6 : 0) + n] + (size_t)shot * b->nf_max * b->fe;
    return IFOS_OK;
}
