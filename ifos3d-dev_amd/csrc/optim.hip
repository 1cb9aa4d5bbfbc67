// Optimiser steps on the device: the L-BFGS two-loop recursion (lbfgs.c),
// its first-iteration gradient save (lbfgs_save.c) and the conjugate-gradient
// direction (conjugategrad.c).  Vector updates are FP32 elementwise in the
// reference's expression order; every dot product is reduced in FP64 (the
// reference sums in FP32 and all-reduces MPI_FLOAT) over a fixed grid of
// partial sums, so a call's result does not depend on scheduling.
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
    hipFree(b->grad);
    hipFree(b->mod);
    hipFree(b->q);
    hipFree(b->part);
    hipFree(b->out);
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
    if (!rc) rc = hip_check(hipDeviceSynchronize(), "lbfgs_savegrad");
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
    if (!rc) rc = hip_check(hipDeviceSynchronize(), "ifos_lbfgs_update");
    return rc;
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
        if (!rc) rc = hip_check(hipDeviceSynchronize(), "ifos_conjugate");
        return rc;
    }
    double *part = nullptr, *out = nullptr;
    int rc = hip_check(hipMalloc(&part, kRedBlocks * 6 * sizeof(double)), "conjugate partials");
    if (!rc) rc = hip_check(hipMalloc(&out, 6 * sizeof(double)), "conjugate sums");
    double s[6] = {0, 0, 0, 0, 0, 0};
    if (!rc) {
        hipLaunchKernelGGL(k_cg_dots, dim3(kRedBlocks), dim3(kRedThreads), 0, 0, G, A, part);
        hipLaunchKernelGGL(k_red_final, dim3(1), dim3(kRedThreads), 0, 0, part, kRedBlocks, 6, out);
        rc = hip_check(hipGetLastError(), "k_cg_dots");
    }
    if (!rc) rc = hip_check(hipMemcpy(s, out, sizeof(s), hipMemcpyDeviceToHost), "conjugate sums copy");
    hipFree(part);
    hipFree(out);
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
    if (!rc) rc = hip_check(hipDeviceSynchronize(), "ifos_conjugate");
    return rc;
}
