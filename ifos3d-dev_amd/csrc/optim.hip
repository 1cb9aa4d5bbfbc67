// Optimiser steps on the device: the L-BFGS two-loop recursion (lbfgs.c),
// its first-iteration gradient save (lbfgs_save.c) and the conjugate-gradient
// direction (conjugategrad.c).  Vector updates are FP32 elementwise in the
// reference's expression order; every dot product is reduced in FP64 (the
// reference sums in FP32 and all-reduces MPI_FLOAT) over a fixed grid of
// partial sums, so a call's result does not depend on scheduling.
#include <algorithm>
#include <cmath>
#include <vector>
#include "ifos_dev.h"

namespace ifos {

constexpr int kRedBlocks = 512;   // partial sums per dot product
constexpr int kRedThreads = 256;

// block-wide FP64 sum of one value per thread (fixed tree order)
__device__ inline double block_sum(double v, double *sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = kRedThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// ---- compact vectors: [numpar][NY][NX][NZ], cell l = ((j-1)*NX + i-1)*NZ + k-1
struct Pack {
    const float *g1, *g2;  // device-layout gradients (g2 used when numpar == 2)
    int numpar;
    int64_t n1;            // NX*NY*NZ
};

__device__ inline int64_t compact_to_field(const DevGeom &g, int64_t l)
{
    const int k = (int)(l % g.nz) + 1;
    const int64_t r = l / g.nz;
    const int i = (int)(r % g.nx) + 1, j = (int)(r / g.nx) + 1;
    return at(g, j, i, k);
}

// mode 0, lbfgs_save.c:31-53: y = grad.  mode 1, lbfgs.c:47-71:
// y += -grad, q = grad.
__global__ __launch_bounds__(256) void k_lb_pack(DevGeom g, Pack P, float *q, float *y, int mode)
{
    const int64_t n = P.n1 * P.numpar;
    for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = compact_to_field(g, l % P.n1);
        const float v = l < P.n1 ? P.g1[c] : P.g2[c];
        if (mode == 0) {
            y[l] = +v;
        } else {
            y[l] += -v;
            q[l] = v;
        }
    }
}

// lbfgs.c:157-178: bfgsgrad[w] = grad, grad = q
__global__ __launch_bounds__(256) void k_lb_unpack(DevGeom g, float *g1, float *g2, int numpar, int64_t n1,
                                                  const float *q, float *y)
{
    const int64_t n = n1 * numpar;
    for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = compact_to_field(g, l % n1);
        float *G = l < n1 ? g1 : g2;
        y[l] = G[c];
        G[c] = q[l];
    }
}

// part[b*2 + 0] = sum a*b, part[b*2 + 1] = sum c*d (when c) over this
// block's share of the compact vectors
__global__ __launch_bounds__(kRedThreads) void k_lb_dot2(const float *a, const float *b, const float *c,
                                                         const float *d, int64_t n, double *part)
{
    __shared__ double sh[kRedThreads];
    double s0 = 0.0, s1 = 0.0;
    for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (int64_t)gridDim.x * blockDim.x) {
        s0 += (double)a[l] * (double)b[l];
        if (c) s1 += (double)c[l] * (double)d[l];
    }
    s0 = block_sum(s0, sh);
    s1 = block_sum(s1, sh);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s0;
        part[2 * blockIdx.x + 1] = s1;
    }
}

// out[m] = sum over blocks of part[b*nd + m], m < nd (one workgroup)
__global__ __launch_bounds__(kRedThreads) void k_red_final(const double *part, int nblk, int nd, double *out)
{
    __shared__ double sh[kRedThreads];
    for (int m = 0; m < nd; ++m) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += part[b * nd + m];
        s = block_sum(s, sh);
        if (threadIdx.x == 0) out[m] = s;
    }
}

// q += a * y (the FP32 update of lbfgs.c:115 and :149 with the coefficient
// already formed); a_first: q += y * a as lbfgs.c:149 orders it
__global__ __launch_bounds__(256) void k_lb_axpy(float *q, const float *y, float a, int64_t n, int a_first)
{
    for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (int64_t)gridDim.x * blockDim.x)
        q[l] = a_first ? q[l] + (a * y[l]) : q[l] + (y[l] * a);
}

// lbfgs.c:120-124: q = q * h0
__global__ __launch_bounds__(256) void k_lb_scale(float *q, float h0, int64_t n)
{
    for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (int64_t)gridDim.x * blockDim.x)
        q[l] = q[l] * h0;
}

// ---- conjugategrad.c on the device-layout fields: rows (j,i) over blocks,
// k over threads
struct CgArgs {
    float *g[3], *p[6];
};

__device__ inline bool cg_row(const DevGeom &g, int64_t r, int &j, int &i)
{
    if (r >= (int64_t)g.nx * g.ny) return false;
    j = (int)(r / g.nx) + 1;
    i = (int)(r % g.nx) + 1;
    return true;
}

// conjugategrad.c:42-54 (cdf == 1): all six priors = grad
__global__ __launch_bounds__(256) void k_cg_copy(DevGeom g, CgArgs A)
{
    for (int64_t r = blockIdx.x; r < (int64_t)g.nx * g.ny; r += gridDim.x) {
        int j, i;
        cg_row(g, r, j, i);
        for (int k = 1 + (int)threadIdx.x; k <= g.nz; k += blockDim.x) {
            const int64_t c = at(g, j, i, k);
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                A.p[n][c] = A.g[n][c];
                A.p[n + 3][c] = A.g[n][c];
            }
        }
    }
}

// conjugategrad.c:58-73: the six sums of the beta ratios, and prior1..3 = grad
__global__ __launch_bounds__(kRedThreads) void k_cg_dots(DevGeom g, CgArgs A, double *part)
{
    __shared__ double sh[kRedThreads];
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t r = blockIdx.x; r < (int64_t)g.nx * g.ny; r += gridDim.x) {
        int j, i;
        cg_row(g, r, j, i);
        for (int k = 1 + (int)threadIdx.x; k <= g.nz; k += blockDim.x) {
            const int64_t c = at(g, j, i, k);
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                const double gn = A.g[n][c], pn = A.p[n][c];
                s[2 * n] += gn * (double)(A.g[n][c] - A.p[n][c]);
                s[2 * n + 1] += pn * pn;
                A.p[n][c] = A.g[n][c];
            }
        }
    }
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        const double t = block_sum(s[m], sh);
        if (threadIdx.x == 0) part[6 * blockIdx.x + m] = t;
    }
}

// conjugategrad.c:93-106: grad = prior4..6 * beta + grad; prior4..6 = grad
__global__ __launch_bounds__(256) void k_cg_update(DevGeom g, CgArgs A, float b0, float b1, float b2)
{
    const float b[3] = {b0, b1, b2};
    for (int64_t r = blockIdx.x; r < (int64_t)g.nx * g.ny; r += gridDim.x) {
        int j, i;
        cg_row(g, r, j, i);
        for (int k = 1 + (int)threadIdx.x; k <= g.nz; k += blockDim.x) {
            const int64_t c = at(g, j, i, k);
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                const float v = (A.p[n + 3][c] * b[n]) + A.g[n][c];
                A.g[n][c] = v;
                A.p[n + 3][c] = v;
            }
        }
    }
}

static unsigned grid_for(int64_t n)
{
    const int64_t b = (n + 255) / 256;
    return (unsigned)(b < 4096 ? (b < 1 ? 1 : b) : 4096);
}

}  // namespace ifos

using namespace ifos;

struct ifos_lbfgs {
    ifos_geom g;
    DevGeom G;
    int numpar, bfgsnum;
    int64_t n1, n;            // cells, numpar * cells
    float *grad = nullptr;    // bfgsgrad [bfgsnum][n] (slot w at w-1)
    float *mod = nullptr;     // bfgsmod  [bfgsnum][n]
    float *q = nullptr;       // [n]
    double *part = nullptr;   // [kRedBlocks * 2]
    double *out = nullptr;    // [2]
    std::vector<double> scale;  // bfgsscale[1..bfgsnum] (index w-1)
};

static int lb_slot(const ifos_lbfgs *b, int v)
{
    int w = v % b->bfgsnum;
    if (w == 0) w = b->bfgsnum;
    return w - 1;
}

// two dot products of compact vectors, reduced in FP64: (a.b, c.d)
static int lb_dot2(ifos_lbfgs *b, const float *x0, const float *y0, const float *x1, const float *y1, double r[2])
{
    hipLaunchKernelGGL(k_lb_dot2, dim3(kRedBlocks), dim3(kRedThreads), 0, 0, x0, y0, x1, y1, b->n, b->part);
    hipLaunchKernelGGL(k_red_final, dim3(1), dim3(kRedThreads), 0, 0, b->part, kRedBlocks, 2, b->out);
    int rc = hip_check(hipGetLastError(), "lbfgs dot");
    if (!rc) rc = hip_check(hipMemcpy(r, b->out, 2 * sizeof(double), hipMemcpyDeviceToHost), "lbfgs dot copy");
    return rc;
}

extern "C" int ifos_lbfgs_create(const ifos_geom *g, int numpar, int bfgsnum, ifos_lbfgs **out)
{
    if (!g || !out || (numpar != 1 && numpar != 2) || bfgsnum < 1) {
        set_error("ifos_lbfgs_create: numpar must be 1 or 2 and bfgsnum >= 1");
        return IFOS_EINVAL;
    }
    *out = nullptr;
    ifos_lbfgs *b = new ifos_lbfgs;
    b->g = *g;
    b->G = make_dev_geom(g);
    b->numpar = numpar;
    b->bfgsnum = bfgsnum;
    b->n1 = (int64_t)g->nx * g->ny * g->nz;
    b->n = b->n1 * numpar;
    b->scale.assign(bfgsnum, 0.0);
    const size_t vb = (size_t)b->n * sizeof(float);
    int rc = hip_check(hipMalloc(&b->grad, vb * bfgsnum), "lbfgs grad");
    if (!rc) rc = hip_check(hipMalloc(&b->mod, vb * bfgsnum), "lbfgs mod");
    if (!rc) rc = hip_check(hipMalloc(&b->q, vb), "lbfgs q");
    if (!rc) rc = hip_check(hipMalloc(&b->part, kRedBlocks * 6 * sizeof(double)), "lbfgs partials");
    if (!rc) rc = hip_check(hipMalloc(&b->out, 6 * sizeof(double)), "lbfgs sums");
    // ifos3d.c:373-379: fmatrix / vector zero-fill
    if (!rc) rc = hip_check(hipMemset(b->grad, 0, vb * bfgsnum), "lbfgs zero");
    if (!rc) rc = hip_check(hipMemset(b->mod, 0, vb * bfgsnum), "lbfgs zero");
    if (rc) {
        ifos_lbfgs_destroy(b);
        return rc;
    }
    *out = b;
    return IFOS_OK;
}

extern "C" void ifos_lbfgs_destroy(ifos_lbfgs *b)
{
    if (!b) return;
    (void)hipFree(b->grad);
    (void)hipFree(b->mod);
    (void)hipFree(b->q);
    (void)hipFree(b->part);
    (void)hipFree(b->out);
    delete b;
}

extern "C" float *ifos_lbfgs_mod(ifos_lbfgs *b, int w)
{
    if (!b || w < 1 || w > b->bfgsnum) return nullptr;
    return b->mod + (size_t)(w - 1) * b->n;
}

extern "C" float *ifos_lbfgs_grad(ifos_lbfgs *b, int w)
{
    if (!b || w < 1 || w > b->bfgsnum) return nullptr;
    return b->grad + (size_t)(w - 1) * b->n;
}

extern "C" int ifos_lbfgs_scale(const ifos_lbfgs *b, double *scale)
{
    if (!b || !scale) return IFOS_EINVAL;
    for (int w = 0; w < b->bfgsnum; ++w) scale[w] = b->scale[w];
    return IFOS_OK;
}

extern "C" int ifos_lbfgs_savegrad(ifos_lbfgs *b, const float *d_grad1, const float *d_grad2)
{
    if (!b || !d_grad1 || (b->numpar == 2 && !d_grad2)) return IFOS_EINVAL;
    const Pack P = {d_grad1, d_grad2, b->numpar, b->n1};
    hipLaunchKernelGGL(k_lb_pack, dim3(grid_for(b->n)), dim3(256), 0, 0, b->G, P, nullptr, b->grad, 0);
    int rc = hip_check(hipGetLastError(), "k_lb_pack");
    if (!rc) rc = hip_check(hipStreamSynchronize(0), "lbfgs_savegrad");
    return rc;
}

extern "C" int ifos_lbfgs_update(ifos_lbfgs *b, float *d_grad1, float *d_grad2, int iteration, double *h0_out)
{
    if (!b || !d_grad1 || (b->numpar == 2 && !d_grad2) || iteration < 2) {
        set_error("ifos_lbfgs_update: iteration must be >= 2 (iteration 1 saves the gradient)");
        return IFOS_EINVAL;
    }
    const int64_t n = b->n;
    const unsigned gr = grid_for(n);
    int m = iteration - b->bfgsnum;
    if (m < 1) m = 1;
    // lbfgs.c:45-71: y_w = g_prev - g, q = g
    int w = lb_slot(b, iteration - 1);
    float *yw = b->grad + (size_t)w * n, *sw = b->mod + (size_t)w * n;
    const Pack P = {d_grad1, d_grad2, b->numpar, b->n1};
    hipLaunchKernelGGL(k_lb_pack, dim3(gr), dim3(256), 0, 0, b->G, P, b->q, yw, 1);
    int rc = hip_check(hipGetLastError(), "k_lb_pack");
    // lbfgs.c:74-88: rho_w = 1 / (y.s), h0 = (y.s) / (y.y)
    double r[2];
    if (!rc) rc = lb_dot2(b, yw, sw, yw, yw, r);
    if (rc) return rc;
    b->scale[w] = 1.0 / r[0];
    const double h0 = r[0] / r[1];
    if (h0_out) *h0_out = h0;
    // lbfgs.c:94-116: newest to oldest, alpha_w = rho_w s_w.q, q -= alpha_w y_w
    std::vector<double> alpha(b->bfgsnum, 0.0);
    for (int v = iteration - 1; v >= m && !rc; --v) {
        w = lb_slot(b, v);
        rc = lb_dot2(b, b->mod + (size_t)w * n, b->q, nullptr, nullptr, r);
        if (rc) break;
        alpha[w] = b->scale[w] * r[0];
        hipLaunchKernelGGL(k_lb_axpy, dim3(gr), dim3(256), 0, 0, b->q, b->grad + (size_t)w * n,
                           (float)(-alpha[w]), n, 1);
        rc = hip_check(hipGetLastError(), "k_lb_axpy");
    }
    // lbfgs.c:120-124
    if (!rc) {
        hipLaunchKernelGGL(k_lb_scale, dim3(gr), dim3(256), 0, 0, b->q, (float)h0, n);
        rc = hip_check(hipGetLastError(), "k_lb_scale");
    }
    // lbfgs.c:136-151: oldest to newest, beta = rho_w y_w.q, q += s_w (alpha_w - beta)
    for (int v = m; v <= iteration - 1 && !rc; ++v) {
        w = lb_slot(b, v);
        rc = lb_dot2(b, b->grad + (size_t)w * n, b->q, nullptr, nullptr, r);
        if (rc) break;
        const double beta = b->scale[w] * r[0];
        hipLaunchKernelGGL(k_lb_axpy, dim3(gr), dim3(256), 0, 0, b->q, b->mod + (size_t)w * n,
                           (float)(alpha[w] - beta), n, 0);
        rc = hip_check(hipGetLastError(), "k_lb_axpy");
    }
    // lbfgs.c:154-178: save this gradient in slot iteration, grad = q
    if (!rc) {
        w = lb_slot(b, iteration);
        hipLaunchKernelGGL(k_lb_unpack, dim3(gr), dim3(256), 0, 0, b->G, d_grad1, d_grad2, b->numpar, b->n1, b->q,
                           b->grad + (size_t)w * n);
        rc = hip_check(hipGetLastError(), "k_lb_unpack");
    }
    if (!rc) rc = hip_check(hipStreamSynchronize(0), "ifos_lbfgs_update");
    return rc;
}

// conjugate's FP64 partial sums: allocated once per device on first use and
// kept for the life of the process (calls on one device share them; the
// optimiser's calls are serialised on the null stream)
static int cg_buffers(double **part, double **out)
{
    static double *s_part[64] = {}, *s_out[64] = {};
    int dev = 0;
    int rc = hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    if (dev < 0 || dev >= 64) {
        set_error("ifos_conjugate: device %d out of range", dev);
        return IFOS_EINVAL;
    }
    if (!s_part[dev]) {
        rc = hip_check(hipMalloc(&s_part[dev], kRedBlocks * 6 * sizeof(double)), "conjugate partials");
        if (!rc) rc = hip_check(hipMalloc(&s_out[dev], 6 * sizeof(double)), "conjugate sums");
        if (rc) {
            if (s_part[dev]) (void)hipFree(s_part[dev]);
            s_part[dev] = nullptr;
            s_out[dev] = nullptr;
            return rc;
        }
    }
    *part = s_part[dev];
    *out = s_out[dev];
    return IFOS_OK;
}

extern "C" int ifos_conjugate(const ifos_geom *g, float *const d_grad[3], float *const d_prior[6], float beta[3],
                              int cdf, double *sums)
{
    if (!g || !d_grad || !d_prior || !beta) return IFOS_EINVAL;
    const DevGeom G = make_dev_geom(g);
    CgArgs A;
    for (int n = 0; n < 3; ++n) A.g[n] = d_grad[n];
    for (int n = 0; n < 6; ++n) A.p[n] = d_prior[n];
    const int64_t rows = (int64_t)g->nx * g->ny;
    const unsigned gr = (unsigned)(rows < 4096 ? rows : 4096);
    beta[0] = beta[1] = beta[2] = 0.0f;  // conjugategrad.c:37
    if (cdf == 1) {
        hipLaunchKernelGGL(k_cg_copy, dim3(gr), dim3(256), 0, 0, G, A);
        int rc = hip_check(hipGetLastError(), "k_cg_copy");
        if (!rc) rc = hip_check(hipStreamSynchronize(0), "ifos_conjugate");
        return rc;
    }
    double *part = nullptr, *out = nullptr;
    int rc = cg_buffers(&part, &out);
    double s[6] = {0, 0, 0, 0, 0, 0};
    if (!rc) {
        hipLaunchKernelGGL(k_cg_dots, dim3(kRedBlocks), dim3(kRedThreads), 0, 0, G, A, part);
        hipLaunchKernelGGL(k_red_final, dim3(1), dim3(kRedThreads), 0, 0, part, kRedBlocks, 6, out);
        rc = hip_check(hipGetLastError(), "k_cg_dots");
    }
    if (!rc) rc = hip_check(hipMemcpy(s, out, sizeof(s), hipMemcpyDeviceToHost), "conjugate sums copy");
    if (rc) return rc;
    if (sums)
        for (int m = 0; m < 6; ++m) sums[m] = s[m];
    // conjugategrad.c:78-90, clamps as written there: only beta[0] ends in
    // [0, 0.5] (the later tests read the already-clamped beta[0])
    if (cdf == 0) {
        if (std::fabs(s[1]) > 0.0) beta[0] = (float)(s[0] / s[1]);
        if (std::fabs(s[3]) > 0.0) beta[1] = (float)(s[2] / s[3]);
        if (std::fabs(s[5]) > 0.0) beta[2] = (float)(s[4] / s[5]);
    }
    if (beta[0] < 0) beta[0] = 0.0f;
    if (beta[0] < 0) beta[1] = 0.0f;
    if (beta[0] < 0) beta[2] = 0.0f;
    if (beta[0] > 0.5f) beta[0] = 0.5f;
    if (beta[0] > 0.5f) beta[1] = 0.5f;
    if (beta[0] > 0.5f) beta[2] = 0.5f;
    hipLaunchKernelGGL(k_cg_update, dim3(gr), dim3(256), 0, 0, G, A, beta[0], beta[1], beta[2]);
    rc = hip_check(hipGetLastError(), "k_cg_update");
    if (!rc) rc = hip_check(hipStreamSynchronize(0), "ifos_conjugate");
    return rc;
}

// ---------------------------------------------------------------------------
// precongrad.c:28-336, modelupdate.c:26-160 and steplength.c:26-91.

namespace ifos {

struct PreconArgs {
    float *g[3];
    const int *src;  // [nsrc][3] grid x, y, z of the sources (iround(srcpos / D))
    const int *rec;  // [ntr][3] grid x, y, z of the receivers
    int nsrc, ntr, damptype, free_surf, sh, rh;
    int nxg, nyg, nzg, posx, posy, posz;
};

// 1 / (1 + c * exp(e)) in double, stored to float as the reference does
__device__ inline float taper(double c, double e)
{
    return (float)(1.0 / (1.0 + c * exp(e)));
}

__global__ __launch_bounds__(256) void k_precon(DevGeom g, PreconArgs A)
{
    for (int64_t r = blockIdx.x; r < (int64_t)g.nx * g.ny; r += gridDim.x) {
        const int j = (int)(r / g.nx) + 1, i = (int)(r % g.nx) + 1;
        for (int k = 1 + (int)threadIdx.x; k <= g.nz; k += blockDim.x) {
            const int64_t c = at(g, j, i, k);
            float a = A.g[0][c], b = A.g[1][c], d = A.g[2][c];
            // global indices: set only by the DAMPTYPE 1 and 2 loops
            // (precongrad.c:45-47, :131-133); 0 elsewhere, as written there
            int ii = 0, jj = 0, kk = 0;
            if (A.damptype == 1 || A.damptype == 2) {
                ii = i + A.posx * g.nx;
                jj = j + A.posy * g.ny;
                kk = k + A.posz * g.nz;
            }
            if (A.damptype == 2) {  // precongrad.c:135-150
                for (int l = 0; l < A.nsrc; ++l) {
                    const int sx = A.src[3 * l], sy = A.src[3 * l + 1], sz = A.src[3 * l + 2];
                    const float rr = (float)sqrt((double)((jj - sy) * (jj - sy) + (ii - sx) * (ii - sx) +
                                                          (kk - sz) * (kk - sz)));
                    const float G = taper(1000.0, -0.7 * (double)rr);
                    a = a * G; b = b * G; d = d * G;
                }
            }
            if (A.damptype == 1 || A.damptype == 2) {  // precongrad.c:49-64, :152-166
                for (int n = 0; n < A.ntr; ++n) {
                    const int rx = A.rec[3 * n], ry = A.rec[3 * n + 1], rz = A.rec[3 * n + 2];
                    const float rr = (float)sqrt((double)((jj - ry) * (jj - ry) + (ii - rx) * (ii - rx) +
                                                          (kk - rz) * (kk - rz)));
                    const float G = taper(1000.0, -0.7 * (double)rr * (double)rr);
                    a = a * G; b = b * G; d = d * G;
                }
            }
            float G0, G1;
            if (A.damptype != 3) {  // precongrad.c:65-113, :167-214, :283-330
                G0 = A.free_surf ? 0.0f : taper(2000.0, -0.6 * (double)(float)(jj * 1.0));
                G1 = taper(2000.0, -0.6 * (double)(float)((A.nyg - jj) * 1.0));
            } else {  // precongrad.c:222-236
                float rr = (float)((jj - A.sh) * 1.0);
                G0 = rr < 0 ? 0.0f : taper(1000.0, -0.7 * (double)rr * (double)rr);
                rr = (float)((A.nyg - jj - A.rh) * 1.0);
                G1 = rr < 0 ? 0.0f : taper(1000.0, -0.7 * (double)rr);
            }
            a = a * G0 * G1; b = b * G0 * G1; d = d * G0 * G1;
            // x and z (the DAMPTYPE 3 tests r < 0 never hold: ii, kk >= 0)
            G0 = taper(2000.0, -0.6 * (double)(float)(ii * 1.0));
            G1 = taper(2000.0, -0.6 * (double)(float)((A.nxg - ii) * 1.0));
            a = a * G0 * G1; b = b * G0 * G1; d = d * G0 * G1;
            G0 = taper(2000.0, -0.6 * (double)(float)(kk * 1.0));
            G1 = taper(2000.0, -0.6 * (double)(float)((A.nzg - kk) * 1.0));
            a = a * G0 * G1; b = b * G0 * G1; d = d * G0 * G1;
            A.g[0][c] = a;
            A.g[1][c] = b;
            A.g[2][c] = d;
        }
    }
}

// block maxima of vp, vs, rho (modelupdate.c:59-74) and of |grad| (:86-95)
__global__ __launch_bounds__(kRedThreads) void k_mu_max(DevGeom g, const float *rho, const float *pi, const float *u,
                                                       const float *g1, const float *g2, const float *g3,
                                                       float *part)
{
    __shared__ float sh[6][kRedThreads];
    float m[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t r = blockIdx.x; r < (int64_t)g.nx * g.ny; r += gridDim.x) {
        const int j = (int)(r / g.nx) + 1, i = (int)(r % g.nx) + 1;
        for (int k = 1 + (int)threadIdx.x; k <= g.nz; k += blockDim.x) {
            const int64_t c = at(g, j, i, k);
            const float vp = (float)sqrt((double)(pi[c] / rho[c]));
            const float vs = (float)sqrt((double)(u[c] / rho[c]));
            if (vp > m[0]) m[0] = vp;
            if (vs > m[1]) m[1] = vs;
            if (rho[c] > m[2]) m[2] = rho[c];
            if (fabsf(g1[c]) > m[3]) m[3] = fabsf(g1[c]);
            if (fabsf(g2[c]) > m[4]) m[4] = fabsf(g2[c]);
            if (fabsf(g3[c]) > m[5]) m[5] = fabsf(g3[c]);
        }
    }
    for (int q = 0; q < 6; ++q) sh[q][threadIdx.x] = m[q];
    __syncthreads();
    for (int s = kRedThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int q = 0; q < 6; ++q) sh[q][threadIdx.x] = fmaxf(sh[q][threadIdx.x], sh[q][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

struct MuArgs {
    float *g[3], *rho, *pi, *u, *mod;  // mod: bfgsmod slot (compact) or null
    float step, s1, s2, s3, vp0, vs0, rho0, n1, n2, n3;  // n*: gradient normalisers (1 when none)
    int numpar;
    int64_t ncell;
};

// modelupdate.c:97-157 (the gradient normalisation folded in as a division)
__global__ __launch_bounds__(256) void k_modelupdate(DevGeom g, MuArgs A, int normalise)
{
    for (int64_t r = blockIdx.x; r < (int64_t)g.nx * g.ny; r += gridDim.x) {
        const int j = (int)(r / g.nx) + 1, i = (int)(r % g.nx) + 1;
        for (int k = 1 + (int)threadIdx.x; k <= g.nz; k += blockDim.x) {
            const int64_t c = at(g, j, i, k);
            float gvp = A.g[0][c], gvs = A.g[1][c], grho = A.g[2][c];
            if (normalise) {
                gvp = gvp / A.n1;
                gvs = gvs / A.n2;
                grho = grho / A.n3;
                A.g[0][c] = gvp;
                A.g[1][c] = gvs;
                A.g[2][c] = grho;
            }
            const float rho = A.rho[c], pi = A.pi[c], u = A.u[c];
            const double vp = sqrt((double)(pi / rho)), vs = sqrt((double)(u / rho));
            const float vpnew = (float)(vp + (double)(((A.step * A.s1) * gvp) * A.vp0));
            const float vsnew = (float)(vs + (double)(((A.step * A.s2) * gvs) * A.vs0));
            const float rhonew = rho + (((A.s3 * A.step) * grho) * A.rho0);
            if (A.mod) {
                const int64_t l = ((int64_t)(j - 1) * g.nx + (i - 1)) * g.nz + (k - 1);
                A.mod[l] = (float)(((double)vpnew - vp) / (double)A.vp0);
                if (A.numpar > 1) A.mod[A.ncell + l] = (float)(((double)vsnew - vs) / (double)A.vs0);
            }
            A.u[c] = (rhonew * vsnew) * vsnew;
            A.pi[c] = (rhonew * vpnew) * vpnew;
            A.rho[c] = rhonew;
        }
    }
}

}  // namespace ifos

extern "C" int ifos_precon_grad(const ifos_geom *g, float *const d_grad[3], int damptype, const int *src_xyz, int nsrc,
                                const int *rec_xyz, int ntr, const int nglob[3], const int pos[3], void *stream)
{
    if (!g || !d_grad || damptype < 0 || damptype > 3 || nsrc < 0 || ntr < 0 || (nsrc && !src_xyz) ||
        (ntr && !rec_xyz)) {
        set_error("ifos_precon_grad: bad arguments (DAMPTYPE must be 0..3)");
        return IFOS_EINVAL;
    }
    const hipStream_t st = (hipStream_t)stream;
    const DevGeom G = make_dev_geom(g);
    PreconArgs A;
    for (int n = 0; n < 3; ++n) A.g[n] = d_grad[n];
    A.nsrc = damptype == 2 ? nsrc : 0;
    A.ntr = (damptype == 1 || damptype == 2) ? ntr : 0;
    A.damptype = damptype;
    A.free_surf = g->free_surf;
    A.sh = A.rh = g->fw + 40;  // precongrad.c:38-39
    A.nxg = nglob ? nglob[0] : g->nx;
    A.nyg = nglob ? nglob[1] : g->ny;
    A.nzg = nglob ? nglob[2] : g->nz;
    A.posx = pos ? pos[0] : 0;
    A.posy = pos ? pos[1] : 0;
    A.posz = pos ? pos[2] : 0;
    int *d = nullptr;
    const size_t nb = (size_t)(A.nsrc + A.ntr) * 3 * sizeof(int);
    int rc = IFOS_OK;
    if (nb) {
        rc = hip_check(hipMalloc(&d, nb), "precon positions");
        if (!rc && A.nsrc)
            rc = hip_check(hipMemcpy(d, src_xyz, (size_t)A.nsrc * 3 * sizeof(int), hipMemcpyHostToDevice), "src");
        if (!rc && A.ntr)
            rc = hip_check(hipMemcpy(d + 3 * A.nsrc, rec_xyz, (size_t)A.ntr * 3 * sizeof(int), hipMemcpyHostToDevice),
                           "rec");
    }
    A.src = d;
    A.rec = d ? d + 3 * A.nsrc : nullptr;
    if (!rc) {
        const int64_t rows = (int64_t)g->nx * g->ny;
        hipLaunchKernelGGL(k_precon, dim3((unsigned)(rows < 8192 ? rows : 8192)), dim3(256), 0, st, G, A);
        rc = hip_check(hipGetLastError(), "k_precon");
    }
    if (!rc) rc = hip_check(hipStreamSynchronize(st), "ifos_precon_grad");
    (void)hipFree(d);
    return rc;
}

extern "C" int ifos_modelupdate(const ifos_geom *g, float *const d_grad[3], float *d_rho, float *d_pi, float *d_u,
                                float *d_bfgsmod, int numpar, int lbfgs, float step, const float weight[3], float vp0,
                                float vs0, float rho0, float maxima[3], void *stream)
{
    if (!g || !d_grad || !d_rho || !d_pi || !d_u || !weight || (lbfgs && !d_bfgsmod)) {
        set_error("ifos_modelupdate: bad arguments");
        return IFOS_EINVAL;
    }
    const hipStream_t st = (hipStream_t)stream;
    const DevGeom G = make_dev_geom(g);
    float *part = nullptr;
    float m[6] = {0, 0, 0, 0, 0, 0};
    int rc = hip_check(hipMalloc(&part, kRedBlocks * 6 * sizeof(float)), "modelupdate partials");
    if (!rc) {
        hipLaunchKernelGGL(k_mu_max, dim3(kRedBlocks), dim3(kRedThreads), 0, st, G, d_rho, d_pi, d_u, d_grad[0],
                           d_grad[1], d_grad[2], part);
        rc = hip_check(hipGetLastError(), "k_mu_max");
    }
    std::vector<float> h(kRedBlocks * 6);
    if (!rc) rc = hip_check(hipMemcpyAsync(h.data(), part, h.size() * sizeof(float), hipMemcpyDeviceToHost, st), "max");
    if (!rc) rc = hip_check(hipStreamSynchronize(st), "modelupdate max");
    (void)hipFree(part);
    if (rc) return rc;
    for (int b = 0; b < kRedBlocks; ++b)
        for (int q = 0; q < 6; ++q) m[q] = std::max(m[q], h[6 * b + q]);
    if (maxima)
        for (int q = 0; q < 3; ++q) maxima[q] = m[q];
    MuArgs A;
    for (int n = 0; n < 3; ++n) A.g[n] = d_grad[n];
    A.rho = d_rho;
    A.pi = d_pi;
    A.u = d_u;
    A.mod = lbfgs ? d_bfgsmod : nullptr;
    A.step = step;
    A.s1 = weight[0];
    A.s2 = weight[1];
    A.s3 = weight[2];
    A.vp0 = vp0;
    A.vs0 = vs0;
    A.rho0 = rho0;
    A.numpar = numpar;
    A.ncell = (int64_t)g->nx * g->ny * g->nz;
    // modelupdate.c:84-115: without L-BFGS the gradients are divided by their
    // maxima, when the vp and vs maxima are both nonzero
    int normalise = 0;
    A.n1 = m[3];
    A.n2 = m[4];
    A.n3 = m[5] == 0.0f ? 1.0f : m[5];
    if (!lbfgs && A.n1 > 0.0f && A.n2 > 0.0f) normalise = 1;
    const int64_t rows = (int64_t)g->nx * g->ny;
    hipLaunchKernelGGL(k_modelupdate, dim3((unsigned)(rows < 8192 ? rows : 8192)), dim3(256), 0, st, G, A, normalise);
    rc = hip_check(hipGetLastError(), "k_modelupdate");
    if (!rc) rc = hip_check(hipStreamSynchronize(st), "ifos_modelupdate");
    return rc;
}

// steplength.c:26-91: the parabola through (step[0..2], L2[0..2]) in T.
// The literals are double as in the reference, so T = float reproduces its
// float/double promotions (0.1 * step[1] formed in double, stored to float);
// T = double is the FP64 fit.
namespace ifos {
template <class T>
__host__ __device__ inline void steplength_t(T L2[4], T step[5], int lbfgs, int it_group, T teststep)
{
    T a = ((step[0] - step[1]) * (L2[0] - L2[2]) - (step[0] - step[2]) * (L2[0] - L2[1])) /
          (step[0] * step[2] * (step[0] - step[2]) + step[0] * step[1] * (step[1] - step[0]) +
           step[1] * step[2] * (step[2] - step[1]));
    T b = (L2[0] - L2[1]) / (step[0] - step[1]);
    b = b - a * (step[0] + step[1]);
    const T c = L2[0] - a * step[0] * step[0] - b * step[0];
    step[3] = 0;
    if (a > 0) step[3] = -b / (2 * a);
    if (a < 0) {
        if (L2[0] < L2[1] && L2[0] < L2[2]) step[3] = 0.1 * step[1];
        if (L2[1] < L2[0] && L2[1] < L2[2]) step[3] = step[1];
        if (L2[2] < L2[0] && L2[2] < L2[1]) step[3] = step[2];
    }
    if (step[3] > 2.5 * step[4]) step[3] = 2.5 * step[4];
    if (lbfgs && it_group > 1 && step[3] > 1.0) step[3] = 1.0;
    if (step[3] <= 0) step[3] = 0.1 * step[1];
    if (step[3] < 0.1 * step[1]) step[3] = 0.1 * step[1];
    L2[3] = a * step[3] * step[3] + b * step[3] + c;
    step[4] = step[3] / 2;
    if (step[4] < 0.2 * teststep) step[4] = 0.2 * teststep;
    if (step[4] > teststep) step[4] = teststep;
    if (lbfgs && it_group > 1) step[4] = 0.5;
}

// one thread: the FP64 fit where the FP64 misfits already are
__global__ void k_steplength64(double *L2, double *step, int lbfgs, int it_group, double teststep)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) steplength_t<double>(L2, step, lbfgs, it_group, teststep);
}
}  // namespace ifos

// steplength.c:26-91 on the host, in the reference's float arithmetic
extern "C" int ifos_steplength(float L2[4], float step[5], int lbfgs, int it_group, float teststep)
{
    if (!L2 || !step) return IFOS_EINVAL;
    ifos::steplength_t<float>(L2, step, lbfgs, it_group, teststep);
    return IFOS_OK;
}

extern "C" int ifos_steplength_f64(double *d_L2, double *d_step, int lbfgs, int it_group, double teststep,
                                   void *stream)
{
    if (!d_L2 || !d_step) {
        set_error("ifos_steplength_f64: null argument");
        return IFOS_EINVAL;
    }
    const hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(ifos::k_steplength64, dim3(1), dim3(64), 0, st, d_L2, d_step, lbfgs, it_group, teststep);
    return hip_check(hipGetLastError(), "k_steplength64");
}
