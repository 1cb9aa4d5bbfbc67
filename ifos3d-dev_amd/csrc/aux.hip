// aux.hip -- the per-step companions of the stencils on gfx950: point
// sources and adjoint injection, receiver gather, free surface, the
// on-the-fly DFT, the frequency-domain gradient, material averaging and the
// single-rank halo wrap of the model, the residual, and a fill kernel.
//
// Arithmetic follows the reference expression by expression; the DFT and
// gradient keep the reference's double-precision steps (disc_fourier.c:42-54,
// gradient_F.c:135-139).
#include <stdio.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "ifos_dev.h"

namespace ifos {

// ---------------------------------------------------------------------------
// sources (update_v_ssg.c:474-522, psource.c:26-54)
__global__ void k_add_forces(DevGeom g, float *vx, float *vy, float *vz, const float *rip, const float *rjp,
                             const float *rkp, const int *src, const float *sig, int nsrc, int nt, int ldsig,
                             float alpha_rad, float beta_rad)
{
    // sources applied in order by one thread: two sources may share a cell
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int l = 0; l < nsrc; ++l) {
        const int i = src[4 * l], j = src[4 * l + 1], k = src[4 * l + 2], type = src[4 * l + 3];
        const int64_t c = at(g, j, i, k);
        const float amp = g.dt * sig[(int64_t)l * ldsig + nt] / (g.dx * g.dy * g.dz);
        switch (type) {
        case 2: vx[c] += amp * rip[c]; break;
        case 3: vz[c] += amp * rjp[c]; break;  // vz with rjp, as the reference pairs them
        case 4: vy[c] += amp * rkp[c]; break;  // vy with rkp
        case 5:
            vx[c] = (float)(vx[c] + cos((double)beta_rad) * cos((double)alpha_rad) * amp);
            vy[c] = (float)(vy[c] + sin((double)alpha_rad) * amp);
            vz[c] = (float)(vz[c] + sin((double)beta_rad) * cos((double)alpha_rad) * amp);
            break;
        default: break;
        }
    }
}

// One thread per receiver.  Receivers that share a cell are applied by the
// lowest-numbered one, in receiver order, so each cell sees the adds in the
// reference's sequence (update_v_ssg.c:512-522).  chain[2r] says whether r is
// the first receiver of its cell, chain[2r+1] is the next receiver of the
// same cell (-1 at the end); without a chain the kernel finds both itself.
__global__ __launch_bounds__(256) void k_add_adjoint(DevGeom g, float *vx, float *vy, float *vz, const int *rec,
                                                    const int *chain, const float *adj, int ntr, int nt, int ldadj)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= ntr) return;
    const int ri = rec[3 * r], rj = rec[3 * r + 1], rk = rec[3 * r + 2];
    if (chain) {
        if (!chain[2 * r]) return;
    } else {
        for (int q = 0; q < r; ++q)
            if (rec[3 * q] == ri && rec[3 * q + 1] == rj && rec[3 * q + 2] == rk) return;
    }
    const int64_t comp = (int64_t)ntr * ldadj;
    const int64_t c = at(g, rj, ri, rk);
    float x = vx[c], y = vy[c], z = vz[c];
    for (int q = r; q >= 0 && q < ntr;) {
        const int64_t s = (int64_t)q * ldadj + nt;
        x += adj[s];
        z += adj[2 * comp + s];
        y += adj[comp + s];
        if (chain) {
            q = chain[2 * q + 1];
        } else {
            int nq = -1;
            for (int p = q + 1; p < ntr; ++p)
                if (rec[3 * p] == ri && rec[3 * p + 1] == rj && rec[3 * p + 2] == rk) {
                    nq = p;
                    break;
                }
            q = nq;
        }
    }
    vx[c] = x;
    vy[c] = y;
    vz[c] = z;
}

__global__ void k_psource(DevGeom g, float *sxx, float *syy, float *szz, const int *src, const float *sig,
                          int nsrc, int nt, int ntall, int ldsig)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    float amp = 0.0f;
    const double den = 2.0 * g.dx * g.dy * g.dz;
    for (int l = 0; l < nsrc; ++l) {
        if (src[4 * l + 3] != 1) continue;
        const float *s = sig + (int64_t)l * ldsig;
        const int64_t c = at(g, src[4 * l + 1], src[4 * l], src[4 * l + 2]);
        if (nt == 1) amp = (float)(s[nt + 1] / den);
        if (nt > 1 && nt < ntall) amp = (float)((s[nt + 1] - s[nt - 1]) / den);
        if (nt == ntall) amp = (float)(-s[nt - 1] / den);
        sxx[c] += amp;
        syy[c] += amp;
        szz[c] += amp;
    }
}

// surface_ssg_elastic.c:78-112 (4th order, j = 1)
__global__ void k_surface(DevGeom g, const float *u, const float *pi, float *sxx, float *syy, float *szz,
                          float *sxy, float *syz, const float *vx, const float *vy, const float *vz)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int i = blockIdx.y + 1;
    if (k > g.nz || i > g.nx) return;
    const int64_t c = at(g, 1, i, k), sj = g.sj, si = g.si;
    const float b1 = g.b1, b2 = g.b2;
    syy[c] = 0.0f;
    for (int m = 1; m <= 2; ++m) {
        syy[c - m * sj] = -syy[c + m * sj];
        sxy[c - m * sj] = -sxy[c + (m - 1) * sj];
        syz[c - m * sj] = -syz[c + (m - 1) * sj];
    }
    float vxx = (b1 * (vx[c] - vx[c - si]) + b2 * (vx[c + si] - vx[c - 2 * si])) / g.dx;
    float vyy = (b1 * (vy[c] - vy[c - sj]) + b2 * (vy[c + sj] - vy[c - 2 * sj])) / g.dy;
    float vzz = (b1 * (vz[c] - vz[c - 1]) + b2 * (vz[c + 1] - vz[c - 2])) / g.dz;
    float f = u[c] * 2.0f;
    float gg = pi[c];
    float h = -(g.dt * (gg - f) * (gg - f) * (vxx + vzz) / gg) - (g.dt * (gg - f) * vyy);
    sxx[c] += h;
    szz[c] += h;
}

// seismo_ssg.c:51-53
__global__ void k_seismo(DevGeom g, const float *vx, const float *vy, const float *vz, const int *rec, int ntr,
                         int ins, int ldsec, float *sec)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= ntr) return;
    const int64_t c = at(g, rec[3 * r + 1], rec[3 * r], rec[3 * r + 2]);
    const int64_t comp = (int64_t)ntr * ldsec, s = (int64_t)r * ldsec + ins;
    sec[s] = vx[c];
    sec[comp + s] = vy[c];
    sec[2 * comp + s] = vz[c];
}

// ---------------------------------------------------------------------------
// on-the-fly DFT, disc_fourier.c:42-54.  One thread per cell of [1..N]^3.
struct Trig {
    double c[IFOS_MAX_NF], s[IFOS_MAX_NF];
};
struct FPtrs {
    float *p[6];
};

__global__ __launch_bounds__(256) void k_dft(DevGeom g, const float *vx, const float *vy, const float *vz,
                                            FPtrs F, int nf, int64_t fstride, Trig T)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int i = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int j = blockIdx.z + 1;
    if (k > g.nz || i > g.nx) return;
    const int64_t c = at(g, j, i, k);
    const double x = vx[c], y = vy[c], z = vz[c];
    for (int l = 0; l < nf; ++l) {
        const int64_t o = c + (int64_t)l * fstride;
        const double t1 = T.c[l], t2 = T.s[l];
        F.p[0][o] = (float)(F.p[0][o] + x * t1);
        F.p[1][o] = (float)(F.p[1][o] + y * t1);
        F.p[2][o] = (float)(F.p[2][o] + z * t1);
        F.p[3][o] = (float)(F.p[3][o] + x * t2);
        F.p[4][o] = (float)(F.p[4][o] + y * t2);
        F.p[5][o] = (float)(F.p[5][o] + z * t2);
    }
}

// compensated variant: S, C float; term in double
__device__ inline void kahan_add(float *S, float *Cc, int64_t o, double term)
{
    float s = S[o], cc = Cc[o];
    float y = (float)(term - (double)cc);
    float t = s + y;
    Cc[o] = (t - s) - y;
    S[o] = t;
}

__global__ __launch_bounds__(256) void k_dft_kahan(DevGeom g, const float *vx, const float *vy, const float *vz,
                                                  FPtrs F, FPtrs Cc, int nf, int64_t fstride, Trig T)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int i = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int j = blockIdx.z + 1;
    if (k > g.nz || i > g.nx) return;
    const int64_t c = at(g, j, i, k);
    const double x = vx[c], y = vy[c], z = vz[c];
    for (int l = 0; l < nf; ++l) {
        const int64_t o = c + (int64_t)l * fstride;
        const double t1 = T.c[l], t2 = T.s[l];
        kahan_add(F.p[0], Cc.p[0], o, x * t1);
        kahan_add(F.p[1], Cc.p[1], o, y * t1);
        kahan_add(F.p[2], Cc.p[2], o, z * t1);
        kahan_add(F.p[3], Cc.p[3], o, x * t2);
        kahan_add(F.p[4], Cc.p[4], o, y * t2);
        kahan_add(F.p[5], Cc.p[5], o, z * t2);
    }
}

// Unrolled variants for nf <= 8: every accumulator load of a cell is in
// flight at once.  Buffer access: one 32-bit offset per cell, frequency l at
// scalar offset l * (field bytes).
template <int NF>
__global__ __launch_bounds__(256) void k_dft_nf(DevGeom g, const float *vx, const float *vy, const float *vz,
                                               FPtrs F, uint32_t tot, Trig T)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int i = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int j = blockIdx.z + 1;
    if (k > g.nz || i > g.nx) return;
    const uint32_t vo = (uint32_t)at(g, j, i, k) * 4u;
    const double x = bload(make_rsrc(vx, g.fbytes), vo, 0);
    const double y = bload(make_rsrc(vy, g.fbytes), vo, 0);
    const double z = bload(make_rsrc(vz, g.fbytes), vo, 0);
    rsrc_t r[6];
#pragma unroll
    for (int n = 0; n < 6; ++n) r[n] = make_rsrc(F.p[n], tot);
    float a[6][NF];
#pragma unroll
    for (int l = 0; l < NF; ++l)
#pragma unroll
        for (int n = 0; n < 6; ++n) a[n][l] = bload(r[n], vo, (uint32_t)l * g.fbytes);
#pragma unroll
    for (int l = 0; l < NF; ++l) {
        const double t1 = T.c[l], t2 = T.s[l];
        bstore((float)(a[0][l] + x * t1), r[0], vo, (uint32_t)l * g.fbytes);
        bstore((float)(a[1][l] + y * t1), r[1], vo, (uint32_t)l * g.fbytes);
        bstore((float)(a[2][l] + z * t1), r[2], vo, (uint32_t)l * g.fbytes);
        bstore((float)(a[3][l] + x * t2), r[3], vo, (uint32_t)l * g.fbytes);
        bstore((float)(a[4][l] + y * t2), r[4], vo, (uint32_t)l * g.fbytes);
        bstore((float)(a[5][l] + z * t2), r[5], vo, (uint32_t)l * g.fbytes);
    }
}

template <int NF>
__global__ __launch_bounds__(256) void k_dft_kahan_nf(DevGeom g, const float *vx, const float *vy,
                                                     const float *vz, FPtrs F, FPtrs Cc, uint32_t tot, Trig T)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int i = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int j = blockIdx.z + 1;
    if (k > g.nz || i > g.nx) return;
    const uint32_t vo = (uint32_t)at(g, j, i, k) * 4u;
    const double v[3] = {bload(make_rsrc(vx, g.fbytes), vo, 0), bload(make_rsrc(vy, g.fbytes), vo, 0),
                         bload(make_rsrc(vz, g.fbytes), vo, 0)};
    rsrc_t r[6], rc[6];
#pragma unroll
    for (int n = 0; n < 6; ++n) {
        r[n] = make_rsrc(F.p[n], tot);
        rc[n] = make_rsrc(Cc.p[n], tot);
    }
    float a[6][NF], cc[6][NF];
#pragma unroll
    for (int l = 0; l < NF; ++l)
#pragma unroll
        for (int n = 0; n < 6; ++n) {
            a[n][l] = bload(r[n], vo, (uint32_t)l * g.fbytes);
            cc[n][l] = bload(rc[n], vo, (uint32_t)l * g.fbytes);
        }
#pragma unroll
    for (int l = 0; l < NF; ++l)
#pragma unroll
        for (int n = 0; n < 6; ++n) {
            const double term = v[n % 3] * (n < 3 ? T.c[l] : T.s[l]);
            const float yv = (float)(term - (double)cc[n][l]);
            const float t = a[n][l] + yv;
            bstore((t - a[n][l]) - yv, rc[n], vo, (uint32_t)l * g.fbytes);
            bstore(t, r[n], vo, (uint32_t)l * g.fbytes);
        }
}

// plain streaming copy (the known-size reference kernel for memory counters):
// four 16-byte loads in flight per thread before the stores
typedef float v4f __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy(v4f *__restrict__ dst, const v4f *__restrict__ src, int64_t n4)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x * 4;
    for (int64_t b = ((int64_t)blockIdx.x * 4) * blockDim.x + threadIdx.x; b < n4; b += stride) {
        v4f v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t t = b + (int64_t)u * blockDim.x;
            if (t < n4) v[u] = __builtin_nontemporal_load(src + t);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t t = b + (int64_t)u * blockDim.x;
            if (t < n4) __builtin_nontemporal_store(v[u], dst + t);
        }
    }
}

// ---------------------------------------------------------------------------
// gradient_F.c:70-72: 1 / (4 pi^2 f^2) of each frequency, the double
// product rounded to float before the reciprocal
struct GradFreq {
    float fdummy[IFOS_MAX_NF];
};

// ---------------------------------------------------------------------------
// matcopy.c:31-307 (single rank): three passes, y faces then x then z
__global__ void k_wrap_y(DevGeom g, float *m)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;  // 0..nz+1
    const int i = blockIdx.y;                              // 0..nx+1
    if (k > g.nz + 1 || i > g.nx + 1) return;
    float top = m[at(g, 1, i, k)], bot = m[at(g, g.ny, i, k)];
    m[at(g, g.ny + 1, i, k)] = top;
    m[at(g, 0, i, k)] = bot;
}
__global__ void k_wrap_x(DevGeom g, float *m)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (k > g.nz + 1 || j > g.ny + 1) return;
    float lef = m[at(g, j, 1, k)], rig = m[at(g, j, g.nx, k)];
    m[at(g, j, g.nx + 1, k)] = lef;
    m[at(g, j, 0, k)] = rig;
}
__global__ void k_wrap_z(DevGeom g, float *m)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i > g.nx + 1 || j > g.ny + 1) return;
    float fro = m[at(g, j, i, 1)], bac = m[at(g, j, i, g.nz)];
    m[at(g, j, i, g.nz + 1)] = fro;
    m[at(g, j, i, 0)] = bac;
}

// av_mat.c:46-64
__global__ __launch_bounds__(256) void k_av_mat(DevGeom g, const float *rho, const float *u, float *uipjp,
                                               float *ujpkp, float *uipkp, float *rjp, float *rkp, float *rip)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int i = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int j = blockIdx.z + 1;
    if (k > g.nz || i > g.nx) return;
    const int64_t c = at(g, j, i, k), sj = g.sj, si = g.si;
    const float *uc = u + c;
    uipjp[c] = (float)(4.0 / ((1.0 / uc[0]) + (1.0 / uc[si]) + (1.0 / uc[sj + si]) + (1.0 / uc[sj])));
    ujpkp[c] = (float)(4.0 / ((1.0 / uc[0]) + (1.0 / uc[1]) + (1.0 / uc[sj + 1]) + (1.0 / uc[sj])));
    uipkp[c] = (float)(4.0 / ((1.0 / uc[0]) + (1.0 / uc[1]) + (1.0 / uc[si + 1]) + (1.0 / uc[si])));
    rjp[c] = (float)(0.5 * (rho[c] + rho[c + sj]));
    rkp[c] = (float)(0.5 * (rho[c] + rho[c + 1]));
    rip[c] = (float)(0.5 * (rho[c] + rho[c + si]));
}

// av_mat.c:56-60 (L > 0): arithmetic averages of taus
__global__ __launch_bounds__(256) void k_av_tau(DevGeom g, const float *taus, float *tausipjp, float *tausjpkp,
                                               float *tausipkp)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int i = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int j = blockIdx.z + 1;
    if (k > g.nz || i > g.nx) return;
    const int64_t c = at(g, j, i, k), sj = g.sj, si = g.si;
    const float *tc = taus + c;
    tausipjp[c] = (float)(0.25 * (tc[0] + tc[si] + tc[sj + si] + tc[sj]));
    tausjpkp[c] = (float)(0.25 * (tc[0] + tc[sj] + tc[sj + 1] + tc[1]));
    tausipkp[c] = (float)(0.25 * (tc[0] + tc[si] + tc[si + 1] + tc[1]));
}

// residual.c:39-53: one thread per trace and component
__global__ void k_residual(const float *obs, const float *syn, float *adj, int ntr, int ns, int ldin, int ldadj,
                           float dt, double *l2part)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;  // comp*ntr + trace
    if (t >= 3 * ntr) return;
    const float *o = obs + (int64_t)t * ldin, *s = syn + (int64_t)t * ldin;
    float *d = adj + (int64_t)t * ldadj;
    float vdiff = 0.0f;
    double acc = 0.0;
    for (int jj = 1; jj <= ns; ++jj) {
        vdiff += s[jj] - o[jj];
        const float dd = dt * vdiff;
        d[ns - jj + 1] = dd;
        acc += (double)(dd * dd);
    }
    l2part[t] = acc;
}

__global__ void k_fill(float *d, int64_t n, float v)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; t < n; t += stride) d[t] = v;
}

// time-domain gradient: running time integrals u += v dt of three
// components over the whole device layout (halos stay 0: v is 0 there)
__global__ void k_td_integrate(float *u0, float *u1, float *u2, const float *v0, const float *v1, const float *v2,
                               int64_t n, float dt)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; t < n; t += stride) {
        u0[t] = u0[t] + v0[t] * dt;
        u1[t] = u1[t] + v1[t] * dt;
        u2[t] = u2[t] + v2[t] * dt;
    }
}

// time-domain gradient: the lambda, mu, rho sums to the vp, vs, rho
// gradients with the parametrisation of gradient_F.c:135-139, in FP64, added
// to the FP32 gradients
__global__ __launch_bounds__(256) void k_td_finish(DevGeom g, const double *al, const double *am, const double *ar,
                                                   const float *rho, const float *pi, const float *u, float *g1,
                                                   float *g2, float *g3)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int i = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int j = blockIdx.z + 1;
    if (k > g.nz || i > g.nx) return;
    const int64_t c = at(g, j, i, k);
    const double r = rho[c], p = pi[c], m = u[c];
    const double srp = sqrt(r * p), sru = sqrt(r * m);
    const double l = al[c], mu = am[c];
    g1[c] = (float)((double)g1[c] + 2.0 * srp * l);
    g2[c] = (float)((double)g2[c] + (-4.0 * sru * l + 2.0 * sru * mu));
    g3[c] = (float)((double)g3[c] + (ar[c] + m / r * mu + (p - 2.0 * m) / r * l));
}

static dim3 cell_grid(const ifos_geom *g, dim3 &block)
{
    block = dim3(64, 4, 1);
    return dim3((unsigned)((g->nz + 63) / 64), (unsigned)((g->nx + 3) / 4), (unsigned)g->ny);
}

}  // namespace ifos

using namespace ifos;

extern "C" int ifos_add_forces(const ifos_geom *g, const ifos_fields *f, const ifos_model *m, const int *d_src,
                               const float *d_sig, int nsrc, int nt, float alpha, float beta, void *stream)
{
    if (nsrc <= 0) return IFOS_OK;
    DevGeom G = make_dev_geom(g);
    const float ar = (float)(alpha * 3.141592653589793 / 180), br = (float)(beta * 3.141592653589793 / 180);
    { ProfScope ps_(IFOS_K_SOURCES, (hipStream_t)stream);
    hipLaunchKernelGGL(k_add_forces, dim3(1), dim3(1), 0, (hipStream_t)stream, G, f->vx, f->vy, f->vz, m->rip,
                       m->rjp, m->rkp, d_src, d_sig, nsrc, nt, g->nt + 2, ar, br);
    }
    return hip_check(hipGetLastError(), "k_add_forces");
}

extern "C" int ifos_add_adjoint(const ifos_geom *g, const ifos_fields *f, const int *d_rec, const int *d_chain,
                                const float *d_adj, int ntr, int nt, void *stream)
{
    if (ntr <= 0) return IFOS_OK;
    DevGeom G = make_dev_geom(g);
    { ProfScope ps_(IFOS_K_SOURCES, (hipStream_t)stream);
    hipLaunchKernelGGL(k_add_adjoint, dim3((unsigned)((ntr + 255) / 256)), dim3(256), 0, (hipStream_t)stream, G,
                       f->vx, f->vy, f->vz, d_rec, d_chain, d_adj, ntr, nt, g->nt + 2);
    }
    return hip_check(hipGetLastError(), "k_add_adjoint");
}

extern "C" int ifos_psource(const ifos_geom *g, const ifos_fields *f, const int *d_src, const float *d_sig,
                            int nsrc, int nt, void *stream)
{
    if (nsrc <= 0) return IFOS_OK;
    DevGeom G = make_dev_geom(g);
    { ProfScope ps_(IFOS_K_SOURCES, (hipStream_t)stream);
    hipLaunchKernelGGL(k_psource, dim3(1), dim3(1), 0, (hipStream_t)stream, G, f->sxx, f->syy, f->szz, d_src,
                       d_sig, nsrc, nt, g->nt, g->nt + 2);
    }
    return hip_check(hipGetLastError(), "k_psource");
}

extern "C" int ifos_surface(const ifos_geom *g, const ifos_fields *f, const ifos_model *m, void *stream)
{
    DevGeom G = make_dev_geom(g);
    { ProfScope ps_(IFOS_K_OTHER, (hipStream_t)stream);
    hipLaunchKernelGGL(k_surface, dim3((unsigned)((g->nz + 63) / 64), (unsigned)g->nx), dim3(64), 0,
                       (hipStream_t)stream, G, m->u, m->pi, f->sxx, f->syy, f->szz, f->sxy, f->syz, f->vx, f->vy,
                       f->vz);
    }
    return hip_check(hipGetLastError(), "k_surface");
}

extern "C" int ifos_seismo(const ifos_geom *g, const ifos_fields *f, const int *d_rec, int ntr, int ins, int ns,
                           float *d_sec, void *stream)
{
    if (ntr <= 0) return IFOS_OK;
    DevGeom G = make_dev_geom(g);
    { ProfScope ps_(IFOS_K_SOURCES, (hipStream_t)stream);
    hipLaunchKernelGGL(k_seismo, dim3((unsigned)((ntr + 255) / 256)), dim3(256), 0, (hipStream_t)stream, G, f->vx,
                       f->vy, f->vz, d_rec, ntr, ins, ns + 1, d_sec);
    }
    return hip_check(hipGetLastError(), "k_seismo");
}

extern "C" int ifos_discfourier(const ifos_geom *g, const ifos_fields *f, float *const d_F[6],
                                float *const d_C[6], const float *finv, int nf, int nt, int ntast, int back,
                                int precision, void *stream)
{
    if (nf <= 0 || nf > IFOS_MAX_NF) {
        set_error("ifos_discfourier: nf=%d outside [1,%d]", nf, IFOS_MAX_NF);
        return IFOS_EINVAL;
    }
    if (precision == IFOS_DFT_KAHAN && !d_C) {
        set_error("ifos_discfourier: Kahan precision needs compensation arrays");
        return IFOS_EINVAL;
    }
    DevGeom G = make_dev_geom(g);
    // time and trig factors exactly as disc_fourier.c:36-44
    float t = 0.0f;
    if (back == 0) t = nt * g->dt;
    if (back == 1) t = g->time - nt * g->dt;
    Trig T;
    for (int l = 0; l < nf; ++l) {
        T.c[l] = cos(2.0 * t * finv[l] * M_PI) * g->dt * ntast;
        T.s[l] = sin(2.0 * t * finv[l] * M_PI) * g->dt * ntast;
    }
    FPtrs F;
    for (int n = 0; n < 6; ++n) F.p[n] = d_F[n];
    dim3 block, grid = cell_grid(g, block);
    const int64_t fs = (int64_t)ifos_field_elems(g);
    const double tot = (double)fs * 4.0 * nf;
    FPtrs Cc = {};
    if (precision == IFOS_DFT_KAHAN)
        for (int n = 0; n < 6; ++n) Cc.p[n] = d_C[n];
    hipStream_t st = (hipStream_t)stream;
    ProfScope ps_(IFOS_K_DFT, st);
    if (nf <= 8 && tot < 4294967295.0) {
        const uint32_t t32 = (uint32_t)tot;
#define IFOS_DFT_CASE(N)                                                                                         \
    case N:                                                                                                      \
        if (precision == IFOS_DFT_KAHAN)                                                                         \
            hipLaunchKernelGGL(k_dft_kahan_nf<N>, grid, block, 0, st, G, f->vx, f->vy, f->vz, F, Cc, t32, T);    \
        else                                                                                                     \
            hipLaunchKernelGGL(k_dft_nf<N>, grid, block, 0, st, G, f->vx, f->vy, f->vz, F, t32, T);              \
        break;
        switch (nf) {
            IFOS_DFT_CASE(1)
            IFOS_DFT_CASE(2)
            IFOS_DFT_CASE(3)
            IFOS_DFT_CASE(4)
            IFOS_DFT_CASE(5)
            IFOS_DFT_CASE(6)
            IFOS_DFT_CASE(7)
            IFOS_DFT_CASE(8)
        }
#undef IFOS_DFT_CASE
    } else if (precision == IFOS_DFT_KAHAN) {
        hipLaunchKernelGGL(k_dft_kahan, grid, block, 0, st, G, f->vx, f->vy, f->vz, F, Cc, nf, fs, T);
    } else {
        hipLaunchKernelGGL(k_dft, grid, block, 0, st, G, f->vx, f->vy, f->vz, F, nf, fs, T);
    }
    return hip_check(hipGetLastError(), "k_dft");
}

extern "C" int ifos_gradient_F(const ifos_geom *g, float *const d_Ff[6], float *const d_Fb[6],
                               const ifos_model *m, const float *finv, int nf, float *const d_grad[3],
                               void *stream)
{
    if (nf <= 0 || nf > IFOS_MAX_NF) {
        set_error("ifos_gradient_F: nf=%d outside [1,%d]", nf, IFOS_MAX_NF);
        return IFOS_EINVAL;
    }
    DevGeom G = make_dev_geom(g);
    GradFreq GF;
    for (int l = 0; l < nf; ++l) {
        // gradient_F.c:70-72: double product stored to float, then 1/x in float
        float fd = (float)(2.0 * 2.0 * finv[l] * finv[l] * M_PI * M_PI);
        GF.fdummy[l] = 1 / fd;
    }
    (void)G;
    // one march launch per frequency, in order (gradient_F.c:66-139)
    const size_t fe = ifos_field_elems(g);
    ProfScope ps_(IFOS_K_GRADIENT, (hipStream_t)stream);
    for (int l = 0; l < nf; ++l) {
        const float *f[12];
        for (int n = 0; n < 6; ++n) {
            f[n] = d_Ff[n] + (size_t)l * fe;
            f[6 + n] = d_Fb[n] + (size_t)l * fe;
        }
        int rc = launch_grad_march(g, f, GF.fdummy[l], m, d_grad, (hipStream_t)stream);
        if (rc) return rc;
    }
    return IFOS_OK;
}

extern "C" int ifos_hess_F(const ifos_geom *g, float *const d_Ff[6], float *const d_Fb[6], int ntr_hess,
                           const ifos_model *m, const float *finv, int nf, float *const d_hess[3], void *stream)
{
    if (nf <= 0 || nf > IFOS_MAX_NF || ntr_hess < 0) {
        set_error("ifos_hess_F: nf=%d outside [1,%d] or ntr_hess=%d < 0", nf, IFOS_MAX_NF, ntr_hess);
        return IFOS_EINVAL;
    }
    const size_t fe = ifos_field_elems(g);
    ProfScope ps_(IFOS_K_GRADIENT, (hipStream_t)stream);
    // hess_F.c:72-77: receivers outer, frequencies inner, one march per pair
    for (int n = 1; n <= ntr_hess; ++n)
        for (int l = 1; l <= nf; ++l) {
            const float fdummy = (float)(finv[l - 1] * M_PI * 2);
            const float fd = 1 / fdummy;
            const size_t mb = (size_t)(n * nf + l - 1);
            const float *f[12];
            for (int c = 0; c < 6; ++c) {
                f[c] = d_Ff[c] + (size_t)(l - 1) * fe;
                f[6 + c] = d_Fb[c] + mb * fe;
            }
            int rc = launch_grad_march(g, f, fd, m, d_hess, (hipStream_t)stream, true);
            if (rc) return rc;
        }
    return IFOS_OK;
}

namespace ifos {
int launch_gradient_shots(const ifos_geom *g, float *const d_Ff[6], float *const d_Fb[6], int64_t fshot, int ns,
                          const ifos_model *m, const float *finv, int nf, float *const d_grad[3], hipStream_t st)
{
    DevGeom G = make_dev_geom(g);
    GradFreq GF;
    for (int l = 0; l < nf; ++l) {
        float fd = (float)(2.0 * 2.0 * finv[l] * finv[l] * M_PI * M_PI);
        GF.fdummy[l] = 1 / fd;
    }
    (void)G;
    // shots in order, frequencies in order: one march launch each
    const size_t fe = ifos_field_elems(g);
    ProfScope ps_(IFOS_K_GRADIENT, st);
    for (int s = 0; s < ns; ++s)
        for (int l = 0; l < nf; ++l) {
            const float *f[12];
            for (int n = 0; n < 6; ++n) {
                f[n] = d_Ff[n] + (size_t)s * fshot + (size_t)l * fe;
                f[6 + n] = d_Fb[n] + (size_t)s * fshot + (size_t)l * fe;
            }
            int rc = launch_grad_march(g, f, GF.fdummy[l], m, d_grad, st);
            if (rc) return rc;
        }
    return IFOS_OK;
}
}  // namespace ifos

extern "C" int ifos_av_tau(const ifos_geom *g, const float *d_taus, float *d_tausipjp, float *d_tausjpkp,
                           float *d_tausipkp, void *stream)
{
    if (!g || !d_taus || !d_tausipjp || !d_tausjpkp || !d_tausipkp) return IFOS_EINVAL;
    DevGeom G = make_dev_geom(g);
    dim3 block, grid = cell_grid(g, block);
    { ProfScope ps_(IFOS_K_OTHER, (hipStream_t)stream);
    hipLaunchKernelGGL(k_av_tau, grid, block, 0, (hipStream_t)stream, G, d_taus, d_tausipjp, d_tausjpkp,
                       d_tausipkp);
    }
    return hip_check(hipGetLastError(), "k_av_tau");
}

extern "C" int ifos_matcopy_wrap(const ifos_geom *g, float *d_m, void *stream)
{
    DevGeom G = make_dev_geom(g);
    hipStream_t s = (hipStream_t)stream;
    { ProfScope ps_(IFOS_K_OTHER, s);
    hipLaunchKernelGGL(k_wrap_y, dim3((unsigned)((g->nz + 2 + 63) / 64), (unsigned)(g->nx + 2)), dim3(64), 0, s, G,
                       d_m);
    }
    { ProfScope ps_(IFOS_K_OTHER, s);
    hipLaunchKernelGGL(k_wrap_x, dim3((unsigned)((g->nz + 2 + 63) / 64), (unsigned)(g->ny + 2)), dim3(64), 0, s, G,
                       d_m);
    }
    { ProfScope ps_(IFOS_K_OTHER, s);
    hipLaunchKernelGGL(k_wrap_z, dim3((unsigned)((g->nx + 2 + 63) / 64), (unsigned)(g->ny + 2)), dim3(64), 0, s, G,
                       d_m);
    }
    return hip_check(hipGetLastError(), "k_wrap");
}

extern "C" int ifos_av_mat(const ifos_geom *g, const float *d_rho, const float *d_u, float *d_uipjp,
                           float *d_ujpkp, float *d_uipkp, float *d_rjp, float *d_rkp, float *d_rip, void *stream)
{
    DevGeom G = make_dev_geom(g);
    dim3 block, grid = cell_grid(g, block);
    { ProfScope ps_(IFOS_K_OTHER, (hipStream_t)stream);
    hipLaunchKernelGGL(k_av_mat, grid, block, 0, (hipStream_t)stream, G, d_rho, d_u, d_uipjp, d_ujpkp, d_uipkp,
                       d_rjp, d_rkp, d_rip);
    }
    return hip_check(hipGetLastError(), "k_av_mat");
}

namespace ifos {
// largest phase velocity at infinite frequency (checkfd_ssg.c:79-122): per
// cell c = sqrt(u (1.0 + L taus) / rho) and sqrt(pi (1.0 + L taup) / rho)
// formed as the C does (L * tau in float, the rest in double, c stored to
// float); cells with rho == 0 are skipped (the reference prints them and
// keeps the previous c).  Maxima of non-negative floats compare as their bit
// patterns, so one unsigned max per workgroup goes to part[blockIdx.x].
__global__ __launch_bounds__(256) void k_cmax(DevGeom g, const float *rho, const float *pi, const float *u,
                                              const float *taup, const float *taus, int L, unsigned *part)
{
    __shared__ unsigned red[256];
    unsigned m = 0u;
    for (int64_t r = blockIdx.x; r < (int64_t)g.nx * g.ny; r += gridDim.x) {
        const int j = (int)(r / g.nx) + 1, i = (int)(r % g.nx) + 1;
        for (int k = 1 + (int)threadIdx.x; k <= g.nz; k += blockDim.x) {
            const int64_t c = at(g, j, i, k);
            const float rr = rho[c];
            if (rr == 0.0f) continue;
            float cs, cp;
            if (L) {
                cs = (float)sqrt((double)u[c] * (1.0 + (double)((float)L * taus[c])) / (double)rr);
                cp = (float)sqrt((double)pi[c] * (1.0 + (double)((float)L * taup[c])) / (double)rr);
            } else {
                cs = (float)sqrt((double)(u[c] / rr));
                cp = (float)sqrt((double)(pi[c] / rr));
            }
            m = max(m, max(__float_as_uint(cs), __float_as_uint(cp)));
        }
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
}  // namespace ifos

extern "C" int ifos_check_stability(const ifos_geom *g, const float *d_rho, const float *d_pi, const float *d_u,
                                    const float *d_taup, const float *d_taus, int L, float *cmax, float *dt_stab,
                                    void *stream)
{
    if (!g || !d_rho || !d_pi || !d_u || (L && (!d_taup || !d_taus)) || L < 0) {
        set_error("ifos_check_stability: null model array (taup, taus needed when L > 0)");
        return IFOS_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    const int64_t rows = (int64_t)g->nx * g->ny;
    const unsigned blocks = (unsigned)std::min<int64_t>(rows, 1024);
    unsigned *part = nullptr;
    if (hip_check(hipMallocAsync((void **)&part, sizeof(unsigned) * blocks, s), "hipMallocAsync")) return IFOS_EHIP;
    hipLaunchKernelGGL(k_cmax, dim3(blocks), dim3(256), 0, s, make_dev_geom(g), d_rho, d_pi, d_u, d_taup, d_taus,
                       L, part);
    std::vector<unsigned> h(blocks);
    int rc = hip_check(hipGetLastError(), "k_cmax");
    if (!rc) rc = hip_check(hipMemcpyAsync(h.data(), part, sizeof(unsigned) * blocks, hipMemcpyDeviceToHost, s),
                            "hipMemcpyAsync");
    if (!rc) rc = hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    (void)hipFreeAsync(part, s);
    if (rc) return rc;
    unsigned mb = 0u;
    for (unsigned v : h) mb = std::max(mb, v);
    float cm;
    memcpy(&cm, &mb, 4);
    // checkfd_ssg.c:475-478 (FDORDER 4), :492-494, :505
    const float gamma = 1.184614f;
    float dhmin = g->dx;
    if (g->dy < dhmin) dhmin = g->dy;
    if (g->dz < dhmin) dhmin = g->dz;
    const float dtstab = (float)(dhmin / (gamma * sqrt(3.0) * cm));
    if (cmax) *cmax = cm;
    if (dt_stab) *dt_stab = dtstab;
    if (g->dt > dtstab) {
        set_error("the simulation will get unstable, choose a smaller DT: DT = %e s, stability limit %e s "
                  "(cmax %.2f m/s)", (double)g->dt, (double)dtstab, (double)cm);
        return IFOS_EUNSTABLE;
    }
    return IFOS_OK;
}

extern "C" int ifos_residual(const ifos_geom *g, const float *d_obs, const float *d_syn, float *d_adj, int ntr,
                             int ns, double *l2, void *stream)
{
    if (ntr <= 0) return IFOS_OK;
    hipStream_t s = (hipStream_t)stream;
    double *part = nullptr;
    if (hip_check(hipMallocAsync((void **)&part, sizeof(double) * 3 * ntr, s), "hipMallocAsync")) return IFOS_EHIP;
    { ProfScope ps_(IFOS_K_OTHER, s);
    hipLaunchKernelGGL(k_residual, dim3((unsigned)((3 * ntr + 127) / 128)), dim3(128), 0, s, d_obs, d_syn, d_adj,
                       ntr, ns, ns + 1, ns + 2, g->dt, part);
    }
    std::vector<double> h(3 * (size_t)ntr);
    int rc = hip_check(hipMemcpyAsync(h.data(), part, sizeof(double) * 3 * ntr, hipMemcpyDeviceToHost, s),
                       "hipMemcpyAsync");
    if (!rc) rc = hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    (void)hipFreeAsync(part, s);
    if (rc) return rc;
    double acc = 0.0;
    for (double v : h) acc += v;
    if (l2) *l2 += acc;
    return IFOS_OK;
}

// ---------------------------------------------------------------------------
// ifos_residual without device allocation.  The per-trace FP64 partial sums
// and their total live in a file-scope device array: part[0 .. 3 ntr) and the
// total at the fixed slot kNoallocTotal.  k_residual writes every partial that
// k_residual_total then reads, so nothing has to start at zero.  Each device
// holds its own copy for the life of the process; every call on a device
// shares it, so calls that overlap on one device race.
constexpr int kNoallocTotal = 3 * IFOS_RESIDUAL_NOALLOC_MAX_TRACES;
static __device__ __align__(256) double scratch[kNoallocTotal + 1];

namespace {

// the partials added in index order, the order ifos_residual's host loop uses
__global__ void k_residual_total(const double *part, int n, double *total)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double acc = 0.0;
    for (int t = 0; t < n; ++t) acc += part[t];
    *total = acc;
}

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices] = {};

// the device address of this device's `scratch`, looked up once per device
int noalloc_scratch(double **out)
{
    int dev = 0;
    int rc = hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("ifos_residual_noalloc: device %d out of range", dev);
        return IFOS_EINVAL;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        rc = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (rc) return rc;
        g_scratch_dev[dev] = (double *)p;
    }
    *out = g_scratch_dev[dev];
    return IFOS_OK;
}
}  // namespace

extern "C" int ifos_residual_noalloc(const ifos_geom *g, const float *d_obs, const float *d_syn, float *d_adj,
                                     int ntr, int ns, double *l2, void *stream)
{
    if (!g) return IFOS_EINVAL;
    if (ntr > IFOS_RESIDUAL_NOALLOC_MAX_TRACES) {
        set_error("ifos_residual_noalloc: %d traces exceed the scratch maximum of %d", ntr,
                  IFOS_RESIDUAL_NOALLOC_MAX_TRACES);
        return IFOS_EINVAL;
    }
    if (ntr <= 0) return IFOS_OK;
    double *part = nullptr;
    int rc = noalloc_scratch(&part);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_residual, dim3((unsigned)((3 * ntr + 127) / 128)), dim3(128), 0, s, d_obs, d_syn, d_adj,
                       ntr, ns, ns + 1, ns + 2, g->dt, part);
    hipLaunchKernelGGL(k_residual_total, dim3(1), dim3(64), 0, s, (const double *)part, 3 * ntr,
                       part + kNoallocTotal);
    rc = hip_check(hipGetLastError(), "k_residual");
    double total = 0.0;
    if (!rc) rc = hip_check(hipMemcpyAsync(&total, part + kNoallocTotal, sizeof(double), hipMemcpyDeviceToHost, s),
                            "hipMemcpyAsync");
    if (!rc) rc = hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (rc) return rc;
    if (l2) *l2 += total;
    return IFOS_OK;
}

namespace ifos {
int launch_td_integrate(const ifos_geom *g, float *const u[3], const float *const v[3], float dt, hipStream_t s)
{
    const size_t n = ifos_field_elems(g);
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    { ProfScope ps_(IFOS_K_GRADIENT, s);
    hipLaunchKernelGGL(k_td_integrate, dim3(blocks), dim3(256), 0, s, u[0], u[1], u[2], v[0], v[1], v[2], (int64_t)n,
                       dt);
    }
    return hip_check(hipGetLastError(), "k_td_integrate");
}

int launch_td_finish(const ifos_geom *g, double *const acc[3], const ifos_model *m, float *const d_grad[3],
                     hipStream_t s)
{
    DevGeom G = make_dev_geom(g);
    dim3 block, grid = cell_grid(g, block);
    { ProfScope ps_(IFOS_K_GRADIENT, s);
    hipLaunchKernelGGL(k_td_finish, grid, block, 0, s, G, acc[0], acc[1], acc[2], m->rho, m->pi, m->u, d_grad[0],
                       d_grad[1], d_grad[2]);
    }
    return hip_check(hipGetLastError(), "k_td_finish");
}
}  // namespace ifos

extern "C" int ifos_fill(float *d, size_t n, float value, void *stream)
{
    if (n == 0) return IFOS_OK;
    unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    { ProfScope ps_(IFOS_K_OTHER, (hipStream_t)stream);
    hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, (hipStream_t)stream, d, (int64_t)n, value);
    }
    return hip_check(hipGetLastError(), "k_fill");
}


extern "C" int ifos_rec_chain(const int *rec_ijk, int ntr, int *chain)
{
    if (ntr < 0 || (ntr > 0 && (!rec_ijk || !chain))) return IFOS_EINVAL;
    for (int r = 0; r < ntr; ++r) {
        chain[2 * r] = 1;
        chain[2 * r + 1] = -1;
    }
    for (int r = 0; r < ntr; ++r) {
        if (!chain[2 * r]) continue;
        int last = r;
        for (int q = r + 1; q < ntr; ++q)
            if (rec_ijk[3 * q] == rec_ijk[3 * r] && rec_ijk[3 * q + 1] == rec_ijk[3 * r + 1] &&
                rec_ijk[3 * q + 2] == rec_ijk[3 * r + 2]) {
                chain[2 * q] = 0;
                chain[2 * last + 1] = q;
                last = q;
            }
    }
    return IFOS_OK;
}

extern "C" int ifos_copy(float *d_dst, const float *d_src, size_t n, void *stream)
{
    if (n == 0) return IFOS_OK;
    if (n % 4 || ((uintptr_t)d_dst | (uintptr_t)d_src) % 16) {
        set_error("ifos_copy: n must be a multiple of 4 and pointers 16-byte aligned");
        return IFOS_EINVAL;
    }
    const int64_t n4 = (int64_t)(n / 4);
    const unsigned blocks = (unsigned)std::min<int64_t>((n4 + 1023) / 1024, 256 * 32);
    ProfScope ps_(IFOS_K_OTHER, (hipStream_t)stream);
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, (hipStream_t)stream, (v4f *)d_dst,
                       (const v4f *)d_src, n4);
    return hip_check(hipGetLastError(), "k_copy");
}
