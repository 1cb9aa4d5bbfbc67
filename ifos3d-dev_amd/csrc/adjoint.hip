// adjoint.hip -- the consistent (exact discrete) adjoint of the elastic
// time step, for gfx950.
//
// No counterpart in the reference, whose adjoint re-runs the forward scheme
// with time-reversed, integrated residuals (residual.c:43-53) and whose
// gradient is the frequency-domain estimate of gradient_F.c.  Here the
// misfit is J = 1/2 sum_n sum_r (syn - obs)^2 of the forward scheme itself
// (update_v + update_v_CPML, update_s_elastic + update_s_CPML_elastic,
// psource, seismo -- the kernels of the plain path) and the gradient is its
// exact derivative: each step's transpose, run backwards.
//
// One reverse step n (forward step n was: v(n) = v(n-1) + d'(sigma(n-1)) /
// rho-bar with the stress-derivative CPML memory, sigma(n) = sigma(n-1) +
// M(model) e'(v(n)) with the velocity-derivative CPML memory, psource,
// seismo(v(n))):
//   1. k_ca_inject: lambda_v += syn(n) - obs(n) at the receivers;
//   2. k_ca_s: per cell, the adjoint of e' (M^T lambda_sigma), through the
//      CPML recursion (adjoint memory m = mu + g; out = g / K + a m;
//      mu <- b m) into scratch, and the model derivatives of the stress
//      update (pi, u, the three averaged u) summed in FP64 with the
//      forward's corrected derivatives e' of v(n);
//   3. k_ca_s_T: lambda_v += D^T scratch (transposed 4th-order differences);
//   4. k_ca_v, 5. k_ca_v_T: the same for the velocity update (adjoint of d'
//      = lambda_v / rho-bar, the rho-bar derivatives with d' of sigma(n-1)),
//      lambda_sigma += D^T scratch.
// k_ca_fold / k_ca_finish carry the sums through av_mat's averages
// (av_mat.c:51-64) and matcopy's halo copies (matcopy.c, one rank: the
// N+1 halo holds cell 1) to the cells, then to (vp, vs, rho) as
// gradient_F.c:135-139 parametrises them.
#include "ifos_dev.h"

namespace ifos {

// derivative components of the velocity update in the order sxx_x sxy_y
// sxz_z syy_y sxy_x syz_z szz_z sxz_x syz_y (update_v_ssg.c:106-123) and of
// the stress update in the order vxx vxy vxz vyx vyy vyz vzx vzy vzz
// (update_s_ssg_elastic.c:106-114): differentiated field (vx vy vz sxx syy
// szz sxy syz sxz), axis, forward (1) / backward (0) difference, half-grid
// CPML profile, CPML memory array (update_v_ssg_CPML.c:84-130,
// update_s_ssg_CPML_elastic.c:63-115)
__constant__ int c_vq_field[9] = {3, 6, 8, 4, 6, 7, 5, 8, 7};
__constant__ int c_vq_axis[9] = {0, 1, 2, 1, 0, 2, 2, 0, 1};
__constant__ int c_vq_fwd[9] = {1, 0, 0, 1, 0, 0, 1, 0, 0};
__constant__ int c_vq_half[9] = {1, 0, 0, 1, 0, 0, 1, 0, 0};
__constant__ int c_vq_psi[9] = {0, 3, 6, 4, 1, 7, 8, 2, 5};
__constant__ int c_sq_field[9] = {0, 0, 0, 1, 1, 1, 2, 2, 2};
__constant__ int c_sq_axis[9] = {0, 1, 2, 0, 1, 2, 0, 1, 2};
__constant__ int c_sq_fwd[9] = {0, 1, 1, 1, 0, 1, 1, 1, 0};
__constant__ int c_sq_half[9] = {0, 1, 1, 1, 0, 1, 1, 1, 0};
__constant__ int c_sq_psi[9] = {0, 3, 6, 1, 4, 7, 2, 5, 8};

struct AdjS {
    float *L[9];                 // adjoint of vx vy vz sxx syy szz sxy syz sxz
    const float *v[3];           // forward v(n)
    Psi9 pv;                     // forward velocity-derivative memory after step n
    Psi9 mu;                     // its adjoint (updated)
    float *scr[9];
    const float *pi, *u, *uipjp, *ujpkp, *uipkp;
    double *G[5];                // d J / d pi, u, uipjp, ujpkp, uipkp
};
struct AdjV {
    float *L[9];
    const float *s[6];           // forward sigma(n-1): sxx syy szz sxy syz sxz
    Psi9 ps;                     // forward stress-derivative memory after step n
    Psi9 mu;
    float *scr[9];
    const float *rip, *rjp, *rkp;
    double *G[3];                // d J / d rip, rjp, rkp
};

__device__ inline bool cell_of(const DevGeom &g, int &j, int &i, int &k)
{
    k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    i = blockIdx.y * blockDim.y + threadIdx.y + 1;
    j = blockIdx.z + 1;
    return k <= g.nz && i <= g.nx;
}

// CPML profile values and memory slot of component q at (j, i, k); false
// where the component has no memory variable (stress: the z layers take K
// from the y axis at the low side, update_s_ssg_CPML_elastic.c:104-108)
__device__ inline bool q_cpml(const DevGeom &g, const Prof &P, bool stress, int q, int j, int i, int k, float &K,
                              float &a, float &b, int &arr, int64_t &at_)
{
    const int ax = stress ? c_sq_axis[q] : c_vq_axis[q];
    const int half = stress ? c_sq_half[q] : c_vq_half[q];
    const int h = ax == 0 ? slab_h(i, g.nx, g.fw, 1) : ax == 1 ? slab_h(j, g.ny, g.fw, !g.free_surf)
                                                               : slab_h(k, g.nz, g.fw, 1);
    if (!h) return false;
    const int kax = (stress && ax == 2) ? (k <= g.fw ? 1 : 2) : ax;
    K = half ? P.KH(kax, h) : P.K(kax, h);
    a = half ? P.AH(ax, h) : P.A(ax, h);
    b = half ? P.BH(ax, h) : P.B(ax, h);
    arr = stress ? c_sq_psi[q] : c_vq_psi[q];
    at_ = ax == 0 ? psi_x_at(g, j, h, k) : ax == 1 ? psi_y_at(g, h, i, k) : psi_z_at(g, j, i, h);
    return true;
}

// transposed differences at x (L zero outside [1..N]^3):
// DFWD^T L = b1 (L[x-s] - L[x]) + b2 (L[x-2s] - L[x+s]),
// DBWD^T L = b1 (L[x] - L[x+s]) + b2 (L[x-s] - L[x+2s])
__device__ inline float dT(const float *L, int64_t x, int64_t s, int fwd, float b1, float b2)
{
    if (fwd) return b1 * (L[x - s] - L[x]) + b2 * (L[x - 2 * s] - L[x + s]);
    return b1 * (L[x] - L[x + s]) + b2 * (L[x - s] - L[x + 2 * s]);
}

__device__ inline int64_t axis_stride(const DevGeom &g, int ax) { return ax == 0 ? g.si : ax == 1 ? g.sj : 1; }

#define D4F(m1, c0, p1, p2) (b1 * ((p1) - (c0)) + b2 * ((p2) - (m1)))
#define D4B(m2, m1, c0, p1) (b1 * ((c0) - (m1)) + b2 * ((p1) - (m2)))

__global__ __launch_bounds__(256) void k_ca_s(DevGeom g, Prof P, AdjS A)
{
    int j, i, k;
    if (!cell_of(g, j, i, k)) return;
    const int64_t c = at(g, j, i, k), sj = g.sj, si = g.si;
    const float b1 = g.b1, b2 = g.b2, DT = g.dt;
    const float Lxx = A.L[3][c], Lyy = A.L[4][c], Lzz = A.L[5][c], Lxy = A.L[6][c], Lyz = A.L[7][c],
                Lxz = A.L[8][c];
    const float gg = A.pi[c], f2 = 2.0f * A.u[c], sumL = Lxx + Lyy + Lzz;
    float gq[9], e[9];
    gq[0] = DT * (gg * sumL - f2 * (Lyy + Lzz));
    gq[4] = DT * (gg * sumL - f2 * (Lxx + Lzz));
    gq[8] = DT * (gg * sumL - f2 * (Lxx + Lyy));
    gq[1] = gq[3] = A.uipjp[c] * DT * Lxy;
    gq[5] = gq[7] = A.ujpkp[c] * DT * Lyz;
    gq[2] = gq[6] = A.uipkp[c] * DT * Lxz;
    // the forward's derivatives of v(n) (update_s_ssg_elastic.c:106-114)
    const float *vx = A.v[0], *vy = A.v[1], *vz = A.v[2];
    e[0] = D4B(vx[c - 2 * si], vx[c - si], vx[c], vx[c + si]) / g.dx;
    e[1] = D4F(vx[c - sj], vx[c], vx[c + sj], vx[c + 2 * sj]) / g.dy;
    e[2] = D4F(vx[c - 1], vx[c], vx[c + 1], vx[c + 2]) / g.dz;
    e[3] = D4F(vy[c - si], vy[c], vy[c + si], vy[c + 2 * si]) / g.dx;
    e[4] = D4B(vy[c - 2 * sj], vy[c - sj], vy[c], vy[c + sj]) / g.dy;
    e[5] = D4F(vy[c - 1], vy[c], vy[c + 1], vy[c + 2]) / g.dz;
    e[6] = D4F(vz[c - si], vz[c], vz[c + si], vz[c + 2 * si]) / g.dx;
    e[7] = D4F(vz[c - sj], vz[c], vz[c + sj], vz[c + 2 * sj]) / g.dy;
    e[8] = D4B(vz[c - 2], vz[c - 1], vz[c], vz[c + 1]) / g.dz;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        float K, a, b;
        int arr;
        int64_t o;
        if (q_cpml(g, P, true, q, j, i, k, K, a, b, arr, o)) {
            const float m = A.mu.p[arr][o] + gq[q];
            e[q] = e[q] / K + A.pv.p[arr][o];
            A.mu.p[arr][o] = b * m;
            A.scr[q][c] = gq[q] / K + a * m;
        } else {
            A.scr[q][c] = gq[q];
        }
    }
    const double dt = (double)DT;
    A.G[0][c] += dt * ((double)e[0] + e[4] + e[8]) * sumL;
    A.G[1][c] += -2.0 * dt * (((double)e[4] + e[8]) * Lxx + ((double)e[0] + e[8]) * Lyy + ((double)e[0] + e[4]) * Lzz);
    A.G[2][c] += dt * ((double)e[1] + e[3]) * Lxy;
    A.G[3][c] += dt * ((double)e[5] + e[7]) * Lyz;
    A.G[4][c] += dt * ((double)e[2] + e[6]) * Lxz;
}

__global__ __launch_bounds__(256) void k_ca_s_T(DevGeom g, AdjS A)
{
    int j, i, k;
    if (!cell_of(g, j, i, k)) return;
    const int64_t c = at(g, j, i, k);
    const float rh[3] = {1.0f / g.dx, 1.0f / g.dy, 1.0f / g.dz};
    float acc[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int q = 0; q < 9; ++q)
        acc[c_sq_field[q]] += rh[c_sq_axis[q]] * dT(A.scr[q], c, axis_stride(g, c_sq_axis[q]), c_sq_fwd[q], g.b1, g.b2);
#pragma unroll
    for (int n = 0; n < 3; ++n) A.L[n][c] += acc[n];
}

__global__ __launch_bounds__(256) void k_ca_v(DevGeom g, Prof P, AdjV A)
{
    int j, i, k;
    if (!cell_of(g, j, i, k)) return;
    const int64_t c = at(g, j, i, k), sj = g.sj, si = g.si;
    const float b1 = g.b1, b2 = g.b2;
    const float rb[3] = {A.rip[c], A.rjp[c], A.rkp[c]};
    const float Lv[3] = {A.L[0][c], A.L[1][c], A.L[2][c]};
    const float hv[3] = {Lv[0] / rb[0], Lv[1] / rb[1], Lv[2] / rb[2]};
    // the forward's derivatives of sigma(n-1) (update_v_ssg.c:106-123)
    const float *sxx = A.s[0], *syy = A.s[1], *szz = A.s[2], *sxy = A.s[3], *syz = A.s[4], *sxz = A.s[5];
    float d[9];
    d[0] = g.dtdx * D4F(sxx[c - si], sxx[c], sxx[c + si], sxx[c + 2 * si]);
    d[1] = g.dtdy * D4B(sxy[c - 2 * sj], sxy[c - sj], sxy[c], sxy[c + sj]);
    d[2] = g.dtdz * D4B(sxz[c - 2], sxz[c - 1], sxz[c], sxz[c + 1]);
    d[3] = g.dtdy * D4F(syy[c - sj], syy[c], syy[c + sj], syy[c + 2 * sj]);
    d[4] = g.dtdx * D4B(sxy[c - 2 * si], sxy[c - si], sxy[c], sxy[c + si]);
    d[5] = g.dtdz * D4B(syz[c - 2], syz[c - 1], syz[c], syz[c + 1]);
    d[6] = g.dtdz * D4F(szz[c - 1], szz[c], szz[c + 1], szz[c + 2]);
    d[7] = g.dtdx * D4B(sxz[c - 2 * si], sxz[c - si], sxz[c], sxz[c + si]);
    d[8] = g.dtdy * D4B(syz[c - 2 * sj], syz[c - sj], syz[c], syz[c + sj]);
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const float gq = hv[q / 3];
        float K, a, b;
        int arr;
        int64_t o;
        if (q_cpml(g, P, false, q, j, i, k, K, a, b, arr, o)) {
            const float m = A.mu.p[arr][o] + gq;
            d[q] = d[q] / K + A.ps.p[arr][o];
            A.mu.p[arr][o] = b * m;
            A.scr[q][c] = gq / K + a * m;
        } else {
            A.scr[q][c] = gq;
        }
    }
#pragma unroll
    for (int n = 0; n < 3; ++n)
        A.G[n][c] -= (double)Lv[n] * ((double)d[3 * n] + d[3 * n + 1] + d[3 * n + 2]) / ((double)rb[n] * rb[n]);
}

__global__ __launch_bounds__(256) void k_ca_v_T(DevGeom g, AdjV A)
{
    int j, i, k;
    if (!cell_of(g, j, i, k)) return;
    const int64_t c = at(g, j, i, k);
    const float dtd[3] = {g.dtdx, g.dtdy, g.dtdz};
    float acc[9] = {};
#pragma unroll
    for (int q = 0; q < 9; ++q)
        acc[c_vq_field[q]] += dtd[c_vq_axis[q]] * dT(A.scr[q], c, axis_stride(g, c_vq_axis[q]), c_vq_fwd[q], g.b1, g.b2);
#pragma unroll
    for (int n = 3; n < 9; ++n) A.L[n][c] += acc[n];
}

// lambda_v += syn(n) - obs(n) at every receiver, in receiver order (one
// thread: receivers may share a cell)
__global__ void k_ca_inject(DevGeom g, float *Lx, float *Ly, float *Lz, const int *rec, int ntr, const float *res,
                            int ldres, int n)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int64_t comp = (int64_t)ntr * ldres;
    for (int r = 0; r < ntr; ++r) {
        const int *p = rec + 3 * r;
        const int64_t c = at(g, p[1], p[0], p[2]), o = (int64_t)r * ldres + n;
        Lx[c] += res[o];
        Ly[c] += res[comp + o];
        Lz[c] += res[2 * comp + o];
    }
}

// res = syn - obs (samples 1..nt of [3][ntr][nt+1]) and J = 1/2 sum res^2 in
// FP64: one workgroup, a fixed assignment of samples to threads and a fixed
// reduction tree (the same sum on every run)
__global__ __launch_bounds__(256) void k_ca_residual(const float *syn, const float *obs, float *res, int64_t rows,
                                                     int nt, double *J)
{
    __shared__ double part[256];
    double s = 0.0;
    const int64_t ld = nt + 1, total = rows * nt;
    for (int64_t q = threadIdx.x; q < total; q += blockDim.x) {
        const int64_t o = (q / nt) * ld + 1 + q % nt;
        const double d = (double)syn[o] - (double)obs[o];
        res[o] = (float)d;
        s += 0.5 * d * d;
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *J = part[0];
}

struct AdjFold {
    const double *G[8];  // rip rjp rkp pi u uipjp ujpkp uipkp
    const float *u, *uipjp, *ujpkp, *uipkp;
    double *Gr, *Gp, *Gu;  // per cell of [1..N+1]^3
};

// the model derivatives gathered per model cell x of [1..N+1]^3 (x = N+1 is
// matcopy's halo): rho from the three density averages that read it
// (rip = (rho(c) + rho(c+i)) / 2 ...), u from the three harmonic averages
// (d uipjp / d u(x) = uipjp^2 / (4 u(x)^2)) and its own term, pi directly
__global__ __launch_bounds__(256) void k_ca_gather(DevGeom g, AdjFold A)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int i = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int j = blockIdx.z + 1;
    if (k > g.nz + 1 || i > g.nx + 1) return;
    const int64_t x = at(g, j, i, k), si = g.si, sj = g.sj;
    const bool in = j <= g.ny && i <= g.nx && k <= g.nz;
    // cells c of [1..N]^3 only contribute (the G arrays are zero elsewhere)
    const double *G0 = A.G[0], *G1 = A.G[1], *G2 = A.G[2];
    A.Gr[x] = 0.5 * ((G0[x] + G0[x - si]) + (G1[x] + G1[x - sj]) + (G2[x] + G2[x - 1]));
    A.Gp[x] = in ? A.G[3][x] : 0.0;
    const double ux = A.u[x], iu2 = 1.0 / (ux * ux);
    double gu = in ? A.G[4][x] : 0.0;
    const int64_t o0[4] = {0, si, sj + si, sj}, o1[4] = {0, 1, sj + 1, sj}, o2[4] = {0, 1, si + 1, si};
    const float *av[3] = {A.uipjp, A.ujpkp, A.uipkp};
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        const int64_t *o = m == 0 ? o0 : (m == 1 ? o1 : o2);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int64_t c = x - o[t];
            const double w = (double)av[m][c] * (double)av[m][c] * 0.25;
            gu += A.G[5 + m][c] * w * iu2;
        }
    }
    A.Gu[x] = gu;
}

// matcopy's halo copies back to their cells (the N+1 layer of an axis holds
// cell 1 of that axis), then (pi, u, rho) -> (vp, vs, rho) with pi = rho vp^2,
// u = rho vs^2 (gradient_F.c:135-139's parametrisation), added to grad
__global__ __launch_bounds__(256) void k_ca_finish(DevGeom g, const double *Gr, const double *Gp, const double *Gu,
                                                   const float *rho, const float *pi, const float *u, float *g1,
                                                   float *g2, float *g3)
{
    int j, i, k;
    if (!cell_of(g, j, i, k)) return;
    const int jj[2] = {j, g.ny + 1}, ii[2] = {i, g.nx + 1}, kk[2] = {k, g.nz + 1};
    const int nj = j == 1 ? 2 : 1, ni = i == 1 ? 2 : 1, nk = k == 1 ? 2 : 1;
    double r_ = 0.0, p_ = 0.0, u_ = 0.0;
    for (int a = 0; a < nj; ++a)
        for (int b = 0; b < ni; ++b)
            for (int d = 0; d < nk; ++d) {
                const int64_t x = at(g, jj[a], ii[b], kk[d]);
                r_ += Gr[x];
                p_ += Gp[x];
                u_ += Gu[x];
            }
    const int64_t c = at(g, j, i, k);
    const double r = rho[c], p = pi[c], m = u[c];
    g1[c] = (float)((double)g1[c] + 2.0 * sqrt(r * p) * p_);
    g2[c] = (float)((double)g2[c] + 2.0 * sqrt(r * m) * u_);
    g3[c] = (float)((double)g3[c] + r_ + p_ * p / r + u_ * m / r);
}

#undef D4F
#undef D4B

static dim3 cells_grid(const ifos_geom *g, int extra = 0)
{
    return dim3((unsigned)((g->nz + extra + 63) / 64), (unsigned)((g->nx + extra + 3) / 4), (unsigned)(g->ny + extra));
}

int launch_ca_step(const ifos_geom *g, const float *d_prof, const CaStep &S, hipStream_t st)
{
    const DevGeom G = make_dev_geom(g);
    const Prof P = {d_prof, 2 * g->fw + 1};
    const dim3 blk(64, 4), grd = cells_grid(g);
    if (S.ntr > 0) {
        hipLaunchKernelGGL(k_ca_inject, dim3(1), dim3(64), 0, st, G, S.L[0], S.L[1], S.L[2], S.rec, S.ntr, S.res,
                           g->nt + 1, S.n);
    }
    AdjS A;
    for (int q = 0; q < 9; ++q) {
        A.L[q] = S.L[q];
        A.scr[q] = S.scr[q];
        A.pv.p[q] = S.pv[q];
        A.mu.p[q] = S.mu_v[q];
    }
    for (int q = 0; q < 3; ++q) A.v[q] = S.v[q];
    A.pi = S.model[1];
    A.u = S.model[2];
    A.uipjp = S.model[3];
    A.ujpkp = S.model[4];
    A.uipkp = S.model[5];
    for (int q = 0; q < 5; ++q) A.G[q] = S.G[3 + q];
    hipLaunchKernelGGL(k_ca_s, grd, blk, 0, st, G, P, A);
    hipLaunchKernelGGL(k_ca_s_T, grd, blk, 0, st, G, A);
    AdjV V;
    for (int q = 0; q < 9; ++q) {
        V.L[q] = S.L[q];
        V.scr[q] = S.scr[q];
        V.ps.p[q] = S.ps[q];
        V.mu.p[q] = S.mu_s[q];
    }
    for (int q = 0; q < 6; ++q) V.s[q] = S.s[q];
    V.rjp = S.model[6];
    V.rkp = S.model[7];
    V.rip = S.model[8];
    for (int q = 0; q < 3; ++q) V.G[q] = S.G[q];
    hipLaunchKernelGGL(k_ca_v, grd, blk, 0, st, G, P, V);
    hipLaunchKernelGGL(k_ca_v_T, grd, blk, 0, st, G, V);
    return hip_check(hipGetLastError(), "consistent adjoint step");
}

int launch_ca_residual(const float *d_syn, const float *d_obs, float *d_res, int ntr, int nt, double *d_J,
                       hipStream_t st)
{
    hipLaunchKernelGGL(k_ca_residual, dim3(1), dim3(256), 0, st, d_syn, d_obs, d_res, (int64_t)3 * ntr, nt, d_J);
    return hip_check(hipGetLastError(), "k_ca_residual");
}

int launch_ca_finish(const ifos_geom *g, double *const G[8], const ifos_model *m, double *Gr, double *Gp,
                     double *Gu, float *const d_grad[3], hipStream_t st)
{
    const DevGeom D = make_dev_geom(g);
    AdjFold F;
    for (int q = 0; q < 8; ++q) F.G[q] = G[q];
    F.u = m->u;
    F.uipjp = m->uipjp;
    F.ujpkp = m->ujpkp;
    F.uipkp = m->uipkp;
    F.Gr = Gr;
    F.Gp = Gp;
    F.Gu = Gu;
    hipLaunchKernelGGL(k_ca_gather, cells_grid(g, 1), dim3(64, 4), 0, st, D, F);
    hipLaunchKernelGGL(k_ca_finish, cells_grid(g), dim3(64, 4), 0, st, D, Gr, Gp, Gu, m->rho, m->pi, m->u,
                       d_grad[0], d_grad[1], d_grad[2]);
    return hip_check(hipGetLastError(), "consistent adjoint gradient");
}

}  // namespace ifos
