// host.cpp -- host side of libifos3d_hip.so: geometry, error reporting, the
// CPML profiles and source wavelet (computed once per run on the CPU, with
// the reference's float/double promotions), and the plain one-shot driver.
#include <algorithm>
#include <cmath>
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <string>
#include <vector>

#include "ifos_dev.h"

namespace ifos {

static thread_local std::string g_err;

void set_error(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

int hip_check(hipError_t e, const char *what)
{
    if (e == hipSuccess) return IFOS_OK;
    set_error("%s: %s", what, hipGetErrorString(e));
    return IFOS_EHIP;
}

DevGeom make_dev_geom(const ifos_geom *g)
{
    DevGeom d;
    d.nx = g->nx;
    d.ny = g->ny;
    d.nz = g->nz;
    d.fw = g->fw;
    d.free_surf = g->free_surf;
    d.si = g->pitch;
    d.sj = (int64_t)(g->nx + 5) * g->pitch;
    d.base = 2 * d.sj + 2 * d.si + IFOS_KOFF;
    d.dx = g->dx;
    d.dy = g->dy;
    d.dz = g->dz;
    d.dt = g->dt;
    d.fbytes = (uint32_t)(ifos_field_elems(g) * 4);
    // exact reciprocal spacings: each spacing 2^e with 2^-e a normal float
    auto pow2 = [](float h, float *r) {
        int e = 0;
        const double m = std::frexp((double)h, &e);
        *r = 1.0f / h;
        return m == 0.5 && std::isnormal(h) && std::isnormal(*r) && (double)*r * (double)h == 1.0;
    };
    d.p2 = pow2(g->dx, &d.rdx) && pow2(g->dy, &d.rdy) && pow2(g->dz, &d.rdz);
    d.dtdx = g->dt / g->dx;
    d.dtdy = g->dt / g->dy;
    d.dtdz = g->dt / g->dz;
    // update_v_ssg.c:98-100
    d.b1 = (float)(9.0 / 8.0);
    d.b2 = (float)(-1.0 / 24.0);
    if (g->fdcoeff == 2) {
        d.b1 = (float)1.1382;
        d.b2 = (float)(-0.046414);
    }
    return d;
}

}  // namespace ifos

using namespace ifos;

extern "C" const char *ifos_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// device memory for callers without a HIP runtime of their own (the tests and
// benchmark use these, so that one HIP runtime serves the whole process)
extern "C" int ifos_dmalloc(void **p, size_t bytes)
{
    if (!p) return IFOS_EINVAL;
    hipError_t e = hipMalloc(p, bytes ? bytes : 4);
    if (e != hipSuccess) {
        set_error("hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
        return IFOS_ENOMEM;
    }
    return IFOS_OK;
}

extern "C" int ifos_dfree(void *p) { return p ? hip_check(hipFree(p), "hipFree") : IFOS_OK; }

extern "C" int ifos_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    return hip_check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

extern "C" int ifos_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    return hip_check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
}

extern "C" int ifos_memcpy_d2d(void *dst, const void *src, size_t bytes)
{
    return hip_check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice), "hipMemcpy D2D");
}

extern "C" int ifos_dmemset(void *p, int value, size_t bytes)
{
    return hip_check(hipMemset(p, value, bytes), "hipMemset");
}

extern "C" int ifos_device_count(int *n)
{
    if (!n) return IFOS_EINVAL;
    *n = 0;
    hipError_t e = hipGetDeviceCount(n);
    if (e != hipSuccess) *n = 0;
    return IFOS_OK;
}

extern "C" int ifos_device_name(char *name, int len)
{
    hipDeviceProp_t prop;
    int rc = hip_check(hipGetDeviceProperties(&prop, 0), "hipGetDeviceProperties");
    if (rc) return rc;
    snprintf(name, (size_t)len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    return IFOS_OK;
}

extern "C" int ifos_device_sync(void) { return hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize"); }

extern "C" int ifos_geom_init(ifos_geom *g, int nx, int ny, int nz, int fw, float dx, float dy, float dz, float dt,
                              int nt, int fdcoeff, int free_surf)
{
    if (!g) return IFOS_EINVAL;
    if (nx < 2 * fw + 1 || ny < 2 * fw + 1 || nz < 2 * fw + 1 || fw < 1) {
        set_error("ifos_geom_init: grid %dx%dx%d too small for FW=%d", nx, ny, nz, fw);
        return IFOS_EINVAL;
    }
    if (!(dx > 0 && dy > 0 && dz > 0 && dt > 0) || nt < 2) {
        set_error("ifos_geom_init: need positive spacings/dt and NT >= 2");
        return IFOS_EINVAL;
    }
    if (fdcoeff != 1 && fdcoeff != 2) {
        set_error("ifos_geom_init: FDCOEFF must be 1 (Taylor) or 2 (Holberg)");
        return IFOS_EINVAL;
    }
    g->nx = nx;
    g->ny = ny;
    g->nz = nz;
    g->fw = fw;
    g->dx = dx;
    g->dy = dy;
    g->dz = dz;
    g->dt = dt;
    g->nt = nt;
    g->time = (nt - 1) * dt;  // ifos3d.c:177
    g->fdcoeff = fdcoeff;
    g->free_surf = free_surf ? 1 : 0;
    g->pitch = ((nz + IFOS_KOFF + 3) + 31) / 32 * 32;  // k in [-2, nz+2] -> [KOFF-2, KOFF+nz+2]
    if ((double)(ny + 5) * (nx + 5) * g->pitch * 4.0 >= 4294967295.0) {
        set_error("ifos_geom_init: one field of %dx%dx%d exceeds 4 GiB (32-bit buffer offsets)", nx, ny, nz);
        return IFOS_EINVAL;
    }
    return IFOS_OK;
}

extern "C" size_t ifos_field_elems(const ifos_geom *g) { return (size_t)(g->ny + 5) * (g->nx + 5) * g->pitch; }

extern "C" size_t ifos_offset(const ifos_geom *g, int j, int i, int k)
{
    DevGeom d = make_dev_geom(g);
    return (size_t)at(d, j, i, k);
}

extern "C" size_t ifos_psi_elems(const ifos_geom *g, int n)
{
    const size_t f2 = 2 * (size_t)g->fw;
    if (n < 3) return (size_t)g->ny * f2 * g->nz;
    if (n < 6) return f2 * g->nx * g->nz;
    return (size_t)g->ny * g->nx * f2;
}

// ---------------------------------------------------------------------------
// CPML_coeff.c:87-324.  |d| is tested with the integer abs() (the float is
// truncated), the y pass clears a_x where (int)d_y == 0, and the right-hand
// side mirrors the left with the integer-grid profile one cell further out.
// C's libm calls on float arguments promote to double; in C++ the float
// overloads (expf, powf, logf) would be picked instead, so every call here
// goes through these double-only wrappers.
static inline double d_exp(double x) { return ::exp(x); }
static inline double d_pow(double x, double y) { return ::pow(x, y); }
static inline double d_log(double x) { return ::log(x); }

static int trunc_nonzero(float d)
{
    int t = (int)d;
    return abs(t) > 1.0e-6;
}

static void cpml_axis(const ifos_geom *g, float dh, float vppml, float fpml, float *K, float *a, float *b,
                      float *Kh, float *ah, float *bh, std::vector<float> &d, float *alpha, float *alpha_h)
{
    const int fw = g->fw;
    const float npower = 3;
    const float k_max_PML = 2.0;
    const float alpha_max_PML = (float)(2.0 * 3.141592653589793 * (fpml / 2.0));
    const float Rcoef = 0.0008f;
    const float ca = 0.25f, cb = 0.75f, cc = 0.0f;
    std::vector<float> dh_(2 * fw + 2, 0.0f);
    const float d0 = (float)(-(npower + 1) * vppml * d_log(Rcoef) / (2.0 * fw * dh));
    for (int i = 1; i <= fw + 1; i++) {
        K[i] = 1.0f;
        float pos = (float)((fw - i + 1) * dh);
        float pn = pos / (float)(fw * dh);
        d[i] = (float)(d0 * (ca * pn + cb * d_pow(pn, npower) + cc * d_pow(pn, 4)));
        K[i] = (float)(1.0 + (k_max_PML - 1.0) * d_pow(pn, npower));
        alpha[i] = (float)(alpha_max_PML * (1.0 - pn));
        b[i] = (float)d_exp(-(d[i] / K[i] + alpha[i]) * g->dt);
        if (trunc_nonzero(d[i])) a[i] = (float)(d[i] * (b[i] - 1.0) / (K[i] * (d[i] + K[i] * alpha[i])));
        if (i <= fw) {
            const int l = 2 * fw - i + 1;
            pos = (float)((fw - i + 0.5) * dh);
            pn = pos / (float)(fw * dh);
            Kh[i] = 1.0f;
            dh_[i] = (float)(d0 * (ca * pn + cb * d_pow(pn, npower) + cc * d_pow(pn, 4)));
            Kh[i] = (float)(1.0 + (k_max_PML - 1.0) * d_pow(pn, npower));
            alpha_h[i] = (float)(alpha_max_PML * (1.0 - pn));
            bh[i] = (float)d_exp(-(dh_[i] / Kh[i] + alpha_h[i]) * g->dt);
            if (trunc_nonzero(dh_[i]))
                ah[i] = (float)(dh_[i] * (bh[i] - 1.0) / (Kh[i] * (dh_[i] + Kh[i] * alpha_h[i])));
            if (i > 1) {
                K[l + 1] = K[i];
                b[l + 1] = b[i];
                if (trunc_nonzero(d[i])) a[l + 1] = a[i];
            }
            Kh[l] = Kh[i];
            bh[l] = bh[i];
            if (trunc_nonzero(d[i])) ah[l] = ah[i];
        }
    }
}

extern "C" int ifos_cpml_profiles(const ifos_geom *g, float vppml, float fpml, float *prof)
{
    if (!g || !prof) return IFOS_EINVAL;
    const int len = 2 * g->fw + 1;
    memset(prof, 0, sizeof(float) * 18 * (size_t)len);
    const float dh[3] = {g->dx, g->dy, g->dz};
    std::vector<float> d(2 * g->fw + 2), al(2 * g->fw + 2), alh(2 * g->fw + 2);
    for (int ax = 0; ax < 3; ++ax) {
        float *p = prof + 6 * ax * len;
        std::fill(d.begin(), d.end(), 0.0f);
        cpml_axis(g, dh[ax], vppml, fpml, p, p + len, p + 2 * len, p + 3 * len, p + 4 * len, p + 5 * len, d,
                  al.data(), alh.data());
        if (ax == 1)
            for (int i = 1; i <= g->fw + 1; ++i)
                if (!trunc_nonzero(d[i])) prof[1 * len + i] = 0.0f;  // a_x, CPML_coeff.c:216
    }
    return IFOS_OK;
}

// alpha_prime profiles of CPML_coeff.c (written for the caller, not used by
// the kernels): out[6*(2fw+1)] = x, x_half, y, y_half, z, z_half
extern "C" int ifos_cpml_alpha(const ifos_geom *g, float fpml, float *out)
{
    if (!g || !out) return IFOS_EINVAL;
    const int fw = g->fw, len = 2 * fw + 1;
    const float alpha_max_PML = (float)(2.0 * 3.141592653589793 * (fpml / 2.0));
    const float dh[3] = {g->dx, g->dy, g->dz};
    memset(out, 0, sizeof(float) * 6 * (size_t)len);
    for (int ax = 0; ax < 3; ++ax) {
        float *al = out + 2 * ax * len, *alh = out + (2 * ax + 1) * len;
        for (int i = 1; i <= fw + 1; i++) {
            float pos = (float)((fw - i + 1) * dh[ax]);
            float pn = pos / (float)(fw * dh[ax]);
            al[i] = (float)(alpha_max_PML * (1.0 - pn));
            if (i <= fw) {
                pos = (float)((fw - i + 0.5) * dh[ax]);
                pn = pos / (float)(fw * dh[ax]);
                alh[i] = (float)(alpha_max_PML * (1.0 - pn));
            }
        }
    }
    return IFOS_OK;
}

// wavelet.c:47-90
extern "C" int ifos_wavelet(const ifos_geom *g, int shape, float tshift, float fc, float amp, float *sig)
{
    if (!g || !sig) return IFOS_EINVAL;
    if (shape != 1 && shape != 2 && shape != 4 && shape != 5) {
        set_error("ifos_wavelet: source shape %d not supported (1, 2, 4, 5)", shape);
        return IFOS_EINVAL;
    }
    const double PI_ = 3.141592653589793;
    sig[0] = 0.0f;
    sig[g->nt + 1] = 0.0f;
    for (int nt = 1; nt <= g->nt; nt++) {
        float t = (float)nt * g->dt;
        float ts = (float)(1.0 / fc), tau, a = 0.0f;
        switch (shape) {
        case 1:
            tau = (float)(PI_ * (t - 1.5 * ts - tshift) / (ts));
            a = (float)((1.0 - 2.0 * tau * tau) * d_exp(-tau * tau));
            break;
        case 2:
            if ((t < tshift) || (t > (tshift + ts)))
                a = 0.0f;
            else
                a = (float)(sin(2.0 * PI_ * (t - tshift) * fc) - 0.5 * sin(4.0 * PI_ * (t - tshift) * fc));
            break;
        case 4:
            if ((t < tshift) || (t > (tshift + ts)))
                a = 0.0f;
            else
                a = (float)((0.75 * PI_ / ts) * d_pow(::sin(PI_ * (t - tshift) / ts), 3.0));
            break;
        case 5: a = (t < tshift) ? 0.0f : 1.0f; break;
        }
        sig[nt] = a * amp;
    }
    return IFOS_OK;
}

extern "C" int ifos_ntast(float fmax, int tast, float dt)
{
    float dummy = (1 / (fmax * tast * dt));
    int n = (int)dummy;
    return n ? n : 1;
}

// ---------------------------------------------------------------------------
// plain one-shot driver
struct ifos_ctx {
    ifos_geom g;
    int nf_max, ntr_max, prec;
    size_t fe;
    float *fields[9];
    float *psi_s[9], *psi_v[9];
    float *model[9];       // rho pi u uipjp ujpkp uipkp rjp rkp rip
    float *prof;
    float *F[12], *C[12];  // forward Re xyz Im xyz, adjoint ...; compensation
    float *sig;            // [nt+2]
    int *src;              // [4]
    int *rec;              // [ntr][3]
    int *chain;            // [ntr][2] same-cell chain of the receivers
    float *adj;            // [3][ntr][nt+2]
    float finv[IFOS_MAX_NF];
    int nf, tast;
    // viscoelastic (L mechanisms): memory variables (L fields per component),
    // relaxation ratios, their averages
    int L, L_alloc;
    float eta[IFOS_MAX_L];
    float *r[6], *taus, *taup, *tavg[3];
    std::vector<void *> allocs;
    struct TdBufs *td;    // buffers of the time-domain gradient (ifos_ctx_run_shot_td)
    struct AdjBufs *cadj;  // buffers of the consistent adjoint (ifos_ctx_run_shot_adj)
};

static void td_free(ifos_ctx *c);
static void adj_free(ifos_ctx *c);

static int dalloc(ifos_ctx *c, void **p, size_t bytes)
{
    if (bytes == 0) bytes = 4;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        set_error("hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
        return IFOS_ENOMEM;
    }
    c->allocs.push_back(*p);
    e = hipMemset(*p, 0, bytes);
    return hip_check(e, "hipMemset");
}

extern "C" int ifos_ctx_destroy(ifos_ctx *c)
{
    if (!c) return IFOS_OK;
    (void)hipDeviceSynchronize();
    td_free(c);
    adj_free(c);
    for (void *p : c->allocs) (void)hipFree(p);
    delete c;
    return IFOS_OK;
}

extern "C" int ifos_ctx_create(ifos_ctx **out, const ifos_geom *g, int nf_max, int ntr_max, int dft_precision)
{
    if (!out || !g || nf_max < 1 || nf_max > IFOS_MAX_NF || ntr_max < 0) {
        set_error("ifos_ctx_create: bad arguments (nf_max in [1,%d])", IFOS_MAX_NF);
        return IFOS_EINVAL;
    }
    ifos_ctx *c = new ifos_ctx();
    c->td = nullptr;
    c->cadj = nullptr;
    c->g = *g;
    c->nf_max = nf_max;
    c->ntr_max = ntr_max;
    c->prec = dft_precision;
    c->fe = ifos_field_elems(g);
    int rc = 0;
    for (int n = 0; n < 9 && !rc; ++n) rc = dalloc(c, (void **)&c->fields[n], c->fe * 4);
    for (int n = 0; n < 9 && !rc; ++n) rc = dalloc(c, (void **)&c->psi_s[n], ifos_psi_elems(g, n) * 4);
    for (int n = 0; n < 9 && !rc; ++n) rc = dalloc(c, (void **)&c->psi_v[n], ifos_psi_elems(g, n) * 4);
    for (int n = 0; n < 9 && !rc; ++n) rc = dalloc(c, (void **)&c->model[n], c->fe * 4);
    if (!rc) rc = dalloc(c, (void **)&c->prof, 18 * (size_t)(2 * g->fw + 1) * 4);
    for (int n = 0; n < 12 && !rc; ++n) rc = dalloc(c, (void **)&c->F[n], c->fe * 4 * (size_t)nf_max);
    for (int n = 0; n < 12; ++n) c->C[n] = nullptr;
    if (dft_precision == IFOS_DFT_KAHAN)
        for (int n = 0; n < 12 && !rc; ++n) rc = dalloc(c, (void **)&c->C[n], c->fe * 4 * (size_t)nf_max);
    if (!rc) rc = dalloc(c, (void **)&c->sig, (size_t)(g->nt + 2) * 4);
    if (!rc) rc = dalloc(c, (void **)&c->src, 4 * sizeof(int));
    if (!rc) rc = dalloc(c, (void **)&c->rec, 3 * sizeof(int) * (size_t)(ntr_max > 0 ? ntr_max : 1));
    if (!rc) rc = dalloc(c, (void **)&c->chain, 2 * sizeof(int) * (size_t)(ntr_max > 0 ? ntr_max : 1));
    if (!rc) rc = dalloc(c, (void **)&c->adj, 3 * (size_t)(ntr_max > 0 ? ntr_max : 1) * (g->nt + 2) * 4);
    if (rc) {
        ifos_ctx_destroy(c);
        return rc;
    }
    c->nf = 0;
    c->tast = 100;
    c->L = 0;
    c->L_alloc = 0;
    for (int l = 0; l < IFOS_MAX_L; ++l) c->eta[l] = 0.0f;
    for (int n = 0; n < 6; ++n) c->r[n] = nullptr;
    c->taus = c->taup = nullptr;
    for (int n = 0; n < 3; ++n) c->tavg[n] = nullptr;
    *out = c;
    return IFOS_OK;
}

// the viscoelastic update (update_s_ssg.c) instead of the elastic one with
// L relaxation mechanisms, eta[l] = DT / relaxation time of mechanism l
// (the reference has L = 1); L = 0 switches back.  taus/taup in device
// layout (cell values), copied and halo-wrapped here.
extern "C" int ifos_ctx_set_visco_n(ifos_ctx *c, const float *d_taus, const float *d_taup, const float *eta,
                                    int L, void *stream)
{
    if (!c || L < 0 || L > IFOS_MAX_L || (L >= 1 && (!d_taus || !d_taup || !eta)) ||
        (double)L * (double)c->fe * 4.0 >= 4294967296.0) {
        set_error("ifos_ctx_set_visco_n: L must be 0..%d (with taus, taup and L etas; L fields under 4 GiB)",
                  IFOS_MAX_L);
        return IFOS_EINVAL;
    }
    c->L = L;
    for (int l = 0; l < IFOS_MAX_L; ++l) c->eta[l] = l < L ? eta[l] : 0.0f;
    if (!L) return IFOS_OK;
    int rc = 0;
    if (!c->taus) {
        if (!rc) rc = dalloc(c, (void **)&c->taus, c->fe * 4);
        if (!rc) rc = dalloc(c, (void **)&c->taup, c->fe * 4);
        for (int n = 0; n < 3 && !rc; ++n) rc = dalloc(c, (void **)&c->tavg[n], c->fe * 4);
        if (rc) return rc;
    }
    if (c->L_alloc < L) {  // memory variables: L fields per component (the old ones stay with the context)
        for (int n = 0; n < 6 && !rc; ++n) rc = dalloc(c, (void **)&c->r[n], c->fe * 4 * (size_t)L);
        if (rc) return rc;
        c->L_alloc = L;
        td_free(c);  // its checkpoint states hold memory variables of the old count
    }
    hipStream_t s = (hipStream_t)stream;
    rc = hip_check(hipMemcpyAsync(c->taus, d_taus, c->fe * 4, hipMemcpyDeviceToDevice, s), "copy taus");
    if (!rc) rc = hip_check(hipMemcpyAsync(c->taup, d_taup, c->fe * 4, hipMemcpyDeviceToDevice, s), "copy taup");
    if (!rc) rc = ifos_matcopy_wrap(&c->g, c->taus, stream);
    if (!rc) rc = ifos_matcopy_wrap(&c->g, c->taup, stream);
    if (!rc) rc = ifos_av_tau(&c->g, c->taus, c->tavg[0], c->tavg[1], c->tavg[2], stream);
    return rc;
}

// one relaxation mechanism, as update_s_ssg.c (L = 0 or 1)
extern "C" int ifos_ctx_set_visco(ifos_ctx *c, const float *d_taus, const float *d_taup, float eta, int L,
                                  void *stream)
{
    if (!c || L < 0 || L > 1) {
        set_error("ifos_ctx_set_visco: L must be 0 or 1 (ifos_ctx_set_visco_n takes more mechanisms)");
        return IFOS_EINVAL;
    }
    return ifos_ctx_set_visco_n(c, d_taus, d_taup, &eta, L, stream);
}

static ifos_model ctx_model(const ifos_ctx *c)
{
    ifos_model m = {c->model[0], c->model[1], c->model[2], c->model[3], c->model[4],
                    c->model[5], c->model[6], c->model[7], c->model[8]};
    return m;
}

static ifos_fields ctx_fields(const ifos_ctx *c)
{
    ifos_fields f = {c->fields[0], c->fields[1], c->fields[2], c->fields[3], c->fields[4],
                     c->fields[5], c->fields[6], c->fields[7], c->fields[8]};
    return f;
}

extern "C" int ifos_ctx_fields(ifos_ctx *c, ifos_fields *f)
{
    *f = ctx_fields(c);
    return IFOS_OK;
}

extern "C" int ifos_ctx_dft(ifos_ctx *c, int back, float *d_F[6])
{
    for (int n = 0; n < 6; ++n) d_F[n] = c->F[(back ? 6 : 0) + n];
    return IFOS_OK;
}

extern "C" int ifos_ctx_set_model(ifos_ctx *c, const float *d_rho, const float *d_pi, const float *d_u, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    const float *in[3] = {d_rho, d_pi, d_u};
    for (int n = 0; n < 3; ++n) {
        int rc = hip_check(hipMemcpyAsync(c->model[n], in[n], c->fe * 4, hipMemcpyDeviceToDevice, s), "copy model");
        if (rc) return rc;
        rc = ifos_matcopy_wrap(&c->g, c->model[n], stream);
        if (rc) return rc;
    }
    return ifos_av_mat(&c->g, c->model[0], c->model[2], c->model[3], c->model[4], c->model[5], c->model[6],
                       c->model[7], c->model[8], stream);
}

extern "C" int ifos_ctx_set_cpml(ifos_ctx *c, float vppml, float fpml)
{
    std::vector<float> h(18 * (size_t)(2 * c->g.fw + 1));
    int rc = ifos_cpml_profiles(&c->g, vppml, fpml, h.data());
    if (rc) return rc;
    return hip_check(hipMemcpy(c->prof, h.data(), h.size() * 4, hipMemcpyHostToDevice), "upload CPML profiles");
}

extern "C" int ifos_ctx_set_freqs(ifos_ctx *c, const float *finv, int nf, int tast)
{
    if (nf < 1 || nf > c->nf_max) {
        set_error("ifos_ctx_set_freqs: nf=%d outside [1,%d]", nf, c->nf_max);
        return IFOS_EINVAL;
    }
    for (int l = 0; l < nf; ++l) c->finv[l] = finv[l];
    c->nf = nf;
    c->tast = tast;
    return IFOS_OK;
}

// the state one propagation advances: wavefields, CPML memory variables and
// (viscoelastic) memory variables
struct ShotState {
    float *fields[9], *psi_s[9], *psi_v[9], *r[6];
};

static ShotState ctx_state(const ifos_ctx *c)
{
    ShotState S;
    for (int n = 0; n < 9; ++n) {
        S.fields[n] = c->fields[n];
        S.psi_s[n] = c->psi_s[n];
        S.psi_v[n] = c->psi_v[n];
    }
    for (int n = 0; n < 6; ++n) S.r[n] = c->r[n];
    return S;
}

// time step nt of a propagation of state S in the order of ifos3d.c:754-814
// (back = 0: forward with the source) and ifos3d.c:905-960 (back = 1:
// adjoint with the receivers injecting): velocity update with its
// injections, stress update, explosive source, the DFT sample (dft), free
// surface, seismograms (d_syn, forward only)
static int step(ifos_ctx *c, const ShotState &S, int back, const ifos_shot *s, int nt, bool dft, int ntast,
                float *d_syn, hipStream_t st)
{
    const ifos_geom *g = &c->g;
    ifos_fields f = {S.fields[0], S.fields[1], S.fields[2], S.fields[3], S.fields[4],
                     S.fields[5], S.fields[6], S.fields[7], S.fields[8]};
    ifos_model m = ctx_model(c);
    ifos_psi ps, pv;
    for (int n = 0; n < 9; ++n) {
        ps.p[n] = S.psi_s[n];
        pv.p[n] = S.psi_v[n];
    }
    int rc = ifos_update_v(g, IFOS_REGION_INTERIOR, &f, &m, c->prof, &ps, st);
    if (!rc) {
        if (!back)
            rc = ifos_add_forces(g, &f, &m, c->src, c->sig, 1, nt, 0.0f, 0.0f, st);
        else
            rc = ifos_add_adjoint(g, &f, c->rec, c->chain, c->adj, s->ntr, nt, st);
    }
    if (!rc) rc = ifos_update_v(g, IFOS_REGION_CPML, &f, &m, c->prof, &ps, st);
    if (!rc && back) rc = ifos_add_adjoint(g, &f, c->rec, c->chain, c->adj, s->ntr, nt, st);
    if (c->L) {
        ifos_visco V;
        for (int n = 0; n < 6; ++n) V.r[n] = S.r[n];
        V.taus = c->taus;
        V.taup = c->taup;
        V.tausipjp = c->tavg[0];
        V.tausjpkp = c->tavg[1];
        V.tausipkp = c->tavg[2];
        V.eta = c->eta[0];
        for (int l = 1; l < IFOS_MAX_L; ++l) V.eta_more[l - 1] = c->eta[l];
        V.L = c->L;
        if (!rc) rc = ifos_update_s_visco(g, IFOS_REGION_INTERIOR, &f, &m, &V, c->prof, &pv, st);
        if (!rc) rc = ifos_update_s_visco(g, IFOS_REGION_CPML, &f, &m, &V, c->prof, &pv, st);
    } else {
        if (!rc) rc = ifos_update_s(g, IFOS_REGION_INTERIOR, &f, &m, c->prof, &pv, st);
        if (!rc) rc = ifos_update_s(g, IFOS_REGION_CPML, &f, &m, c->prof, &pv, st);
    }
    if (!rc && !back) rc = ifos_psource(g, &f, c->src, c->sig, 1, nt, st);
    if (!rc && dft && nt % ntast == 0) {
        float *Fp[6], *Cp[6];
        for (int n = 0; n < 6; ++n) {
            Fp[n] = c->F[(back ? 6 : 0) + n];
            Cp[n] = c->C[(back ? 6 : 0) + n];
        }
        rc = ifos_discfourier(g, &f, Fp, c->prec == IFOS_DFT_KAHAN ? Cp : nullptr, c->finv, c->nf, nt, ntast, back,
                              c->prec, st);
    }
    if (!rc && g->free_surf) rc = ifos_surface(g, &f, &m, st);
    if (!rc && !back && d_syn) rc = ifos_seismo(g, &f, c->rec, s->ntr, nt, g->nt, d_syn, st);
    return rc;
}

// NaN probe of ifos3d.c:761 on vy of state S (the reference checks every
// step on the host; here every 64 steps)
static int probe_nan(ifos_ctx *c, const ShotState &S, int nt, hipStream_t st)
{
    const ifos_geom *g = &c->g;
    const size_t probe = ifos_offset(g, g->ny / 2, g->nx / 2, g->nz / 2);
    float v = 0.0f;
    int rc = hip_check(hipMemcpyAsync(&v, S.fields[1] + probe, 4, hipMemcpyDeviceToHost, st), "probe");
    if (!rc) rc = hip_check(hipStreamSynchronize(st), "probe sync");
    if (!rc && isnan(v)) {
        set_error("simulation unstable: NaN at the probe cell (step %d)", nt);
        rc = IFOS_ENAN;
    }
    return rc;
}

// one propagation of NT steps; back = 0 forward (sources), 1 adjoint
static int propagate(ifos_ctx *c, int back, const ifos_shot *s, int ntast, float *d_syn, hipStream_t st)
{
    const ShotState S = ctx_state(c);
    const bool dft = s->want_adjoint != 0;
    for (int nt = 1; nt <= c->g.nt; ++nt) {
        int rc = (nt & 63) == 0 ? probe_nan(c, S, nt, st) : IFOS_OK;
        if (!rc) rc = step(c, S, back, s, nt, dft, ntast, d_syn, st);
        if (rc) return rc;
    }
    return IFOS_OK;
}

static int zero_wavefield(ifos_ctx *c, hipStream_t st)
{
    int rc = 0;
    for (int n = 0; n < 9 && !rc; ++n) rc = ifos_fill(c->fields[n], c->fe, 0.0f, st);
    for (int n = 0; n < 9 && !rc; ++n) rc = ifos_fill(c->psi_s[n], ifos_psi_elems(&c->g, n), 0.0f, st);
    for (int n = 0; n < 9 && !rc; ++n) rc = ifos_fill(c->psi_v[n], ifos_psi_elems(&c->g, n), 0.0f, st);
    if (c->L)  // zero_wavefield.c:80-95
        for (int n = 0; n < 6 && !rc; ++n) rc = ifos_fill(c->r[n], c->fe * (size_t)c->L, 0.0f, st);
    return rc;
}

// checks a shot and uploads its wavelet, source and receivers
static int prepare_shot(ifos_ctx *c, const ifos_shot *s, hipStream_t st, const char *who, bool need_freqs)
{
    const ifos_geom *g = &c->g;
    if (s->ntr > c->ntr_max) {
        set_error("%s: ntr=%d > ntr_max=%d", who, s->ntr, c->ntr_max);
        return IFOS_EINVAL;
    }
    if (need_freqs && s->want_adjoint && c->nf < 1) {
        set_error("%s: frequencies not set", who);
        return IFOS_EINVAL;
    }
    // bounds of source and receivers
    if (s->src_i < 1 || s->src_i > g->nx || s->src_j < 1 || s->src_j > g->ny || s->src_k < 1 || s->src_k > g->nz) {
        set_error("%s: source (%d,%d,%d) outside the grid", who, s->src_i, s->src_j, s->src_k);
        return IFOS_EINVAL;
    }
    for (int r = 0; r < s->ntr; ++r) {
        const int *p = s->rec_ijk + 3 * r;
        if (p[0] < 1 || p[0] > g->nx || p[1] < 1 || p[1] > g->ny || p[2] < 1 || p[2] > g->nz) {
            set_error("%s: receiver %d (%d,%d,%d) outside the grid", who, r, p[0], p[1], p[2]);
            return IFOS_EINVAL;
        }
    }
    std::vector<float> sig(g->nt + 2);
    int rc = ifos_wavelet(g, s->source_shape, s->tshift, s->fc, s->amp, sig.data());
    if (rc) return rc;
    int hsrc[4] = {s->src_i, s->src_j, s->src_k, s->src_type};
    rc = hip_check(hipMemcpyAsync(c->sig, sig.data(), sig.size() * 4, hipMemcpyHostToDevice, st), "upload signal");
    if (!rc) rc = hip_check(hipMemcpyAsync(c->src, hsrc, sizeof hsrc, hipMemcpyHostToDevice, st), "upload source");
    std::vector<int> chain(2 * (size_t)(s->ntr > 0 ? s->ntr : 1));
    if (!rc && s->ntr > 0) rc = ifos_rec_chain(s->rec_ijk, s->ntr, chain.data());
    if (!rc && s->ntr > 0)
        rc = hip_check(hipMemcpyAsync(c->rec, s->rec_ijk, 3 * sizeof(int) * s->ntr, hipMemcpyHostToDevice, st),
                       "upload receivers");
    if (!rc && s->ntr > 0)
        rc = hip_check(hipMemcpyAsync(c->chain, chain.data(), 2 * sizeof(int) * s->ntr, hipMemcpyHostToDevice, st),
                       "upload receiver chain");
    if (!rc) rc = hip_check(hipStreamSynchronize(st), "sync uploads");
    return rc;
}

extern "C" int ifos_ctx_run_shot(ifos_ctx *c, const ifos_shot *s, float *d_syn, float *const d_grad[3],
                                 double *l2, void *stream)
{
    if (!c || !s) return IFOS_EINVAL;
    const ifos_geom *g = &c->g;
    hipStream_t st = (hipStream_t)stream;
    int rc = prepare_shot(c, s, st, "ifos_ctx_run_shot", true);
    if (rc) return rc;
    const int ntast = s->want_adjoint ? ifos_ntast(c->finv[c->nf - 1], c->tast, g->dt) : 1;
    // ifos3d.c:738-739: zero wavefield and DFT accumulators
    rc = zero_wavefield(c, st);
    for (int n = 0; n < 12 && !rc; ++n) {
        rc = ifos_fill(c->F[n], c->fe * c->nf_max, 0.0f, st);
        if (!rc && c->C[n]) rc = ifos_fill(c->C[n], c->fe * c->nf_max, 0.0f, st);
    }
    if (!rc && d_syn) rc = ifos_fill(d_syn, 3 * (size_t)s->ntr * (g->nt + 1), 0.0f, st);
    if (!rc) rc = propagate(c, 0, s, ntast, d_syn, st);
    if (rc || !s->want_adjoint) return rc;

    // residual -> adjoint sources (ifos3d.c:837-866), zero adjoint traces first
    rc = ifos_fill(c->adj, 3 * (size_t)s->ntr * (g->nt + 2), 0.0f, st);
    if (!rc && s->d_obs && d_syn) rc = ifos_residual(g, s->d_obs, d_syn, c->adj, s->ntr, g->nt, l2, st);
    if (!rc) rc = zero_wavefield(c, st);
    if (!rc) rc = propagate(c, 1, s, ntast, nullptr, st);
    if (rc) return rc;
    ifos_model m = ctx_model(c);
    float *Ff[6], *Fb[6];
    for (int n = 0; n < 6; ++n) {
        Ff[n] = c->F[n];
        Fb[n] = c->F[6 + n];
    }
    return ifos_gradient_F(g, Ff, Fb, &m, c->finv, c->nf, d_grad, st);
}

// ---------------------------------------------------------------------------
// time-domain gradient with checkpoints (no counterpart in the reference,
// whose gradient is the frequency-domain one of gradient_F.c).  The products
// of gradient_F.c:122-132 are summed over the time samples of the shot
// instead of over DFT frequencies: lambda and mu from the derivatives of the
// displacements (the velocities integrated from the end of the record,
// u(t) = int_t^T v dt', which is where the 1 / (4 pi^2 f^2) of gradient_F
// comes from), rho from the velocities, each sample weighted by DT.  The
// forward sample n pairs with the adjoint step m = NT - 1 - n, the pairing
// of the reference's DFT times (disc_fourier.c:36-37: t = n DT forward,
// TIME - m DT adjoint, TIME = (NT - 1) DT), for n = 1 .. NT - 2.
// The forward pass stores its state every K steps; the adjoint pass runs
// the segments from the last to the first, recomputing each segment's
// forward velocities from its checkpoint (one extra forward propagation).
struct TdBufs {
    int K = 0, nck = 0, L = 0;  // L: memory-variable mechanisms of the states
    std::vector<ShotState> ck;  // state after steps 0, K, 2K, ...
    ShotState work;             // the recomputed forward state
    std::vector<float *> seg;   // [K][3] forward velocities of one segment
    float *uf[3], *ub[3];       // running displacements, forward and adjoint
    double *acc[3];             // lambda, mu, rho sums
    std::vector<void *> allocs;
};

static void td_free(ifos_ctx *c)
{
    if (!c->td) return;
    for (void *p : c->td->allocs) (void)hipFree(p);
    delete c->td;
    c->td = nullptr;
}

static int td_alloc_one(TdBufs *T, void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes ? bytes : 4);
    if (e != hipSuccess) {
        set_error("hipMalloc(%zu) for the time-domain gradient: %s", bytes, hipGetErrorString(e));
        return IFOS_ENOMEM;
    }
    T->allocs.push_back(*p);
    return IFOS_OK;
}

static int td_alloc_state(ifos_ctx *c, TdBufs *T, ShotState &S)
{
    int rc = IFOS_OK;
    for (int n = 0; n < 9 && !rc; ++n) rc = td_alloc_one(T, (void **)&S.fields[n], c->fe * 4);
    for (int n = 0; n < 9 && !rc; ++n) rc = td_alloc_one(T, (void **)&S.psi_s[n], ifos_psi_elems(&c->g, n) * 4);
    for (int n = 0; n < 9 && !rc; ++n) rc = td_alloc_one(T, (void **)&S.psi_v[n], ifos_psi_elems(&c->g, n) * 4);
    for (int n = 0; n < 6; ++n) S.r[n] = nullptr;
    if (c->L)
        for (int n = 0; n < 6 && !rc; ++n) rc = td_alloc_one(T, (void **)&S.r[n], c->fe * 4 * (size_t)c->L);
    return rc;
}

static int td_alloc(ifos_ctx *c, int K)
{
    const int nck = (c->g.nt + K - 1) / K;
    if (c->td && c->td->K == K && c->td->L == c->L) return IFOS_OK;
    td_free(c);
    TdBufs *T = new TdBufs();
    c->td = T;
    T->K = K;
    T->L = c->L;
    T->nck = nck;
    T->ck.resize(nck);
    int rc = IFOS_OK;
    for (int q = 0; q < nck && !rc; ++q) rc = td_alloc_state(c, T, T->ck[q]);
    if (!rc) rc = td_alloc_state(c, T, T->work);
    T->seg.assign(3 * (size_t)K, nullptr);
    for (size_t q = 0; q < T->seg.size() && !rc; ++q) rc = td_alloc_one(T, (void **)&T->seg[q], c->fe * 4);
    for (int n = 0; n < 3 && !rc; ++n) rc = td_alloc_one(T, (void **)&T->uf[n], c->fe * 4);
    for (int n = 0; n < 3 && !rc; ++n) rc = td_alloc_one(T, (void **)&T->ub[n], c->fe * 4);
    for (int n = 0; n < 3 && !rc; ++n) rc = td_alloc_one(T, (void **)&T->acc[n], c->fe * 8);
    if (rc) td_free(c);
    return rc;
}

static int copy_state(ifos_ctx *c, const ShotState &dst, const ShotState &src, hipStream_t st)
{
    int rc = IFOS_OK;
    for (int n = 0; n < 9 && !rc; ++n)
        rc = hip_check(hipMemcpyAsync(dst.fields[n], src.fields[n], c->fe * 4, hipMemcpyDeviceToDevice, st),
                       "checkpoint copy");
    for (int n = 0; n < 9 && !rc; ++n) {
        const size_t b = ifos_psi_elems(&c->g, n) * 4;
        rc = hip_check(hipMemcpyAsync(dst.psi_s[n], src.psi_s[n], b, hipMemcpyDeviceToDevice, st), "checkpoint copy");
        if (!rc)
            rc = hip_check(hipMemcpyAsync(dst.psi_v[n], src.psi_v[n], b, hipMemcpyDeviceToDevice, st),
                           "checkpoint copy");
    }
    if (c->L)
        for (int n = 0; n < 6 && !rc; ++n)
            rc = hip_check(hipMemcpyAsync(dst.r[n], src.r[n], c->fe * 4 * (size_t)c->L, hipMemcpyDeviceToDevice, st),
                           "checkpoint copy");
    return rc;
}

extern "C" size_t ifos_td_bytes(const ifos_geom *g, int every, int L)
{
    const int K = every > 0 ? std::min(every, g->nt) : std::max(1, (int)lround(sqrt((double)g->nt)));
    const size_t nck = (size_t)(g->nt + K - 1) / K, fe = ifos_field_elems(g);
    size_t psi = 0;
    for (int n = 0; n < 9; ++n) psi += ifos_psi_elems(g, n);
    const size_t state = (9 + 6 * (size_t)L) * fe * 4 + 2 * psi * 4;
    return (nck + 1) * state + 3 * (size_t)K * fe * 4 + 6 * fe * 4 + 3 * fe * 8;
}

extern "C" int ifos_ctx_run_shot_td(ifos_ctx *c, const ifos_shot *s, int every, float *d_syn,
                                    float *const d_grad[3], double *l2, void *stream)
{
    if (!c || !s || !d_syn || !d_grad) {
        set_error("ifos_ctx_run_shot_td: needs a shot, the seismogram buffer and the gradients");
        return IFOS_EINVAL;
    }
    const ifos_geom *g = &c->g;
    const int N = g->nt;
    hipStream_t st = (hipStream_t)stream;
    if (!s->want_adjoint || !s->d_obs || N < 3) {
        set_error("ifos_ctx_run_shot_td: needs want_adjoint, observed data and NT >= 3");
        return IFOS_EINVAL;
    }
    const int K = every > 0 ? std::min(every, N) : std::max(1, (int)lround(sqrt((double)N)));
    int rc = prepare_shot(c, s, st, "ifos_ctx_run_shot_td", false);
    if (!rc) rc = td_alloc(c, K);
    if (rc) return rc;
    TdBufs *T = c->td;
    const ShotState S = ctx_state(c);
    // forward (ifos3d.c:754-814) with the state stored after steps 0, K, 2K, ...
    rc = zero_wavefield(c, st);
    if (!rc) rc = ifos_fill(d_syn, 3 * (size_t)s->ntr * (N + 1), 0.0f, st);
    if (!rc) rc = copy_state(c, T->ck[0], S, st);
    for (int nt = 1; nt <= N && !rc; ++nt) {
        if ((nt & 63) == 0) rc = probe_nan(c, S, nt, st);
        if (!rc) rc = step(c, S, 0, s, nt, false, 1, d_syn, st);
        if (!rc && nt % K == 0 && nt / K < T->nck) rc = copy_state(c, T->ck[nt / K], S, st);
    }
    // residual -> adjoint sources (ifos3d.c:837-866)
    if (!rc) rc = ifos_fill(c->adj, 3 * (size_t)s->ntr * (N + 2), 0.0f, st);
    if (!rc) rc = ifos_residual(g, s->d_obs, d_syn, c->adj, s->ntr, N, l2, st);
    if (!rc) rc = zero_wavefield(c, st);
    for (int n = 0; n < 3 && !rc; ++n) {
        rc = ifos_fill(T->uf[n], c->fe, 0.0f, st);
        if (!rc) rc = ifos_fill(T->ub[n], c->fe, 0.0f, st);
        if (!rc) rc = ifos_fill((float *)T->acc[n], 2 * c->fe, 0.0f, st);
    }
    const float dt = g->dt;
    const float *vb[3] = {S.fields[0], S.fields[1], S.fields[2]};
    for (int q = T->nck - 1; q >= 0 && !rc; --q) {
        const int a = q * K, b = std::min(a + K, N);
        // recompute the segment's forward velocities from its checkpoint
        rc = copy_state(c, T->work, T->ck[q], st);
        for (int n = a + 1; n <= b && !rc; ++n) {
            rc = step(c, T->work, 0, s, n, false, 1, nullptr, st);
            for (int d = 0; d < 3 && !rc; ++d)
                rc = hip_check(hipMemcpyAsync(T->seg[3 * (n - a - 1) + d], T->work.fields[d], c->fe * 4,
                                              hipMemcpyDeviceToDevice, st),
                               "segment copy");
        }
        // the segment backwards, the adjoint forwards (ifos3d.c:905-960)
        for (int n = b; n > a && !rc; --n) {
            const float *vf[3] = {T->seg[3 * (n - a - 1)], T->seg[3 * (n - a - 1) + 1], T->seg[3 * (n - a - 1) + 2]};
            rc = launch_td_integrate(g, T->uf, vf, dt, st);
            if (rc || n > N - 2) continue;
            const int m = N - 1 - n;
            if ((m & 63) == 0) rc = probe_nan(c, S, m, st);
            if (!rc) rc = step(c, S, 1, s, m, false, 1, nullptr, st);
            if (!rc) rc = launch_td_integrate(g, T->ub, vb, dt, st);
            if (!rc) rc = launch_td_sample(g, T->uf, vf, T->ub, vb, (double)dt, T->acc, st);
        }
    }
    ifos_model mod = ctx_model(c);
    if (!rc) rc = launch_td_finish(g, T->acc, &mod, d_grad, st);
    return rc;
}

// ---------------------------------------------------------------------------
// consistent (exact discrete) adjoint (adjoint.hip).  The forward pass is
// ifos_ctx_run_shot's, with the state stored every K steps; the reverse pass
// walks the segments from the last to the first, recomputes each segment's
// forward steps from its checkpoint into a K-deep ring of what the
// transposed steps read (v(n), sigma(n-1) and both CPML memories after step
// n), then runs the reverse steps of the segment.
struct AdjBufs {
    int K = 0, nck = 0, ntr = 0;
    std::vector<ShotState> ck;
    ShotState work;
    std::vector<float *> rv, rs, rps, rpv;  // ring: [K][3] v, [K][6] sigma, [K][9] psi_s, [K][9] psi_v
    float *L[9], *scr[9], *mu_s[9], *mu_v[9];
    double *G[8], *Gr, *Gp, *Gu, *J;
    float *res;
    std::vector<void *> allocs;
};

static void adj_free(ifos_ctx *c)
{
    if (!c->cadj) return;
    for (void *p : c->cadj->allocs) (void)hipFree(p);
    delete c->cadj;
    c->cadj = nullptr;
}

static int adj_one(AdjBufs *A, void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes ? bytes : 4);
    if (e != hipSuccess) {
        set_error("hipMalloc(%zu) for the consistent adjoint: %s", bytes, hipGetErrorString(e));
        return IFOS_ENOMEM;
    }
    A->allocs.push_back(*p);
    return hip_check(hipMemset(*p, 0, bytes ? bytes : 4), "hipMemset");
}

static int adj_alloc(ifos_ctx *c, int K, int ntr)
{
    if (c->cadj && c->cadj->K == K && c->cadj->ntr == ntr) return IFOS_OK;
    adj_free(c);
    AdjBufs *A = new AdjBufs();
    c->cadj = A;
    A->K = K;
    A->ntr = ntr;
    A->nck = (c->g.nt + K - 1) / K;
    const size_t fe = c->fe;
    int rc = IFOS_OK;
    auto state = [&](ShotState &S) {
        for (int n = 0; n < 9 && !rc; ++n) rc = adj_one(A, (void **)&S.fields[n], fe * 4);
        for (int n = 0; n < 9 && !rc; ++n) rc = adj_one(A, (void **)&S.psi_s[n], ifos_psi_elems(&c->g, n) * 4);
        for (int n = 0; n < 9 && !rc; ++n) rc = adj_one(A, (void **)&S.psi_v[n], ifos_psi_elems(&c->g, n) * 4);
        for (int n = 0; n < 6; ++n) S.r[n] = nullptr;
    };
    A->ck.resize(A->nck);
    for (int q = 0; q < A->nck && !rc; ++q) state(A->ck[q]);
    if (!rc) state(A->work);
    A->rv.assign(3 * (size_t)K, nullptr);
    A->rs.assign(6 * (size_t)K, nullptr);
    A->rps.assign(9 * (size_t)K, nullptr);
    A->rpv.assign(9 * (size_t)K, nullptr);
    for (auto &p : A->rv)
        if (!rc) rc = adj_one(A, (void **)&p, fe * 4);
    for (auto &p : A->rs)
        if (!rc) rc = adj_one(A, (void **)&p, fe * 4);
    for (size_t q = 0; q < A->rps.size() && !rc; ++q) {
        rc = adj_one(A, (void **)&A->rps[q], ifos_psi_elems(&c->g, (int)(q % 9)) * 4);
        if (!rc) rc = adj_one(A, (void **)&A->rpv[q], ifos_psi_elems(&c->g, (int)(q % 9)) * 4);
    }
    for (int n = 0; n < 9 && !rc; ++n) {
        rc = adj_one(A, (void **)&A->L[n], fe * 4);
        if (!rc) rc = adj_one(A, (void **)&A->scr[n], fe * 4);
        if (!rc) rc = adj_one(A, (void **)&A->mu_s[n], ifos_psi_elems(&c->g, n) * 4);
        if (!rc) rc = adj_one(A, (void **)&A->mu_v[n], ifos_psi_elems(&c->g, n) * 4);
    }
    for (int n = 0; n < 8 && !rc; ++n) rc = adj_one(A, (void **)&A->G[n], fe * 8);
    if (!rc) rc = adj_one(A, (void **)&A->Gr, fe * 8);
    if (!rc) rc = adj_one(A, (void **)&A->Gp, fe * 8);
    if (!rc) rc = adj_one(A, (void **)&A->Gu, fe * 8);
    if (!rc) rc = adj_one(A, (void **)&A->J, 8);
    if (!rc) rc = adj_one(A, (void **)&A->res, 3 * (size_t)(ntr > 0 ? ntr : 1) * (c->g.nt + 1) * 4);
    if (rc) adj_free(c);
    return rc;
}

static int adj_K(const ifos_geom *g, int every)
{
    return every > 0 ? std::min(every, g->nt) : std::max(1, (int)lround(sqrt((double)g->nt)));
}

extern "C" size_t ifos_adj_bytes(const ifos_geom *g, int every, int ntr)
{
    const int K = adj_K(g, every);
    const size_t nck = (size_t)(g->nt + K - 1) / K, fe = ifos_field_elems(g);
    size_t psi = 0;
    for (int n = 0; n < 9; ++n) psi += ifos_psi_elems(g, n);
    const size_t state = 9 * fe * 4 + 2 * psi * 4;
    return (nck + 1) * state + (size_t)K * state + 18 * fe * 4 + 2 * psi * 4 + 11 * fe * 8 +
           3 * (size_t)(ntr > 0 ? ntr : 1) * (g->nt + 1) * 4 + 8;
}

extern "C" int ifos_ctx_run_shot_adj(ifos_ctx *c, const ifos_shot *s, int every, float *d_syn,
                                     float *const d_grad[3], double *J, void *stream)
{
    if (!c || !s || !d_syn || !d_grad || !s->d_obs) {
        set_error("ifos_ctx_run_shot_adj: needs a shot with observed data, the seismogram buffer and the gradients");
        return IFOS_EINVAL;
    }
    const ifos_geom *g = &c->g;
    if (c->L || g->free_surf || s->src_type != 1) {
        set_error("ifos_ctx_run_shot_adj: the elastic update, no free surface and an explosive source only");
        return IFOS_EINVAL;
    }
    const int N = g->nt, K = adj_K(g, every);
    hipStream_t st = (hipStream_t)stream;
    int rc = prepare_shot(c, s, st, "ifos_ctx_run_shot_adj", false);
    if (!rc) rc = adj_alloc(c, K, s->ntr);
    if (rc) return rc;
    AdjBufs *A = c->cadj;
    const ShotState S = ctx_state(c);
    // forward (ifos3d.c:754-814) with the state stored after steps 0, K, 2K, ...
    rc = zero_wavefield(c, st);
    if (!rc) rc = ifos_fill(d_syn, 3 * (size_t)s->ntr * (N + 1), 0.0f, st);
    if (!rc) rc = copy_state(c, A->ck[0], S, st);
    for (int nt = 1; nt <= N && !rc; ++nt) {
        if ((nt & 63) == 0) rc = probe_nan(c, S, nt, st);
        if (!rc) rc = step(c, S, 0, s, nt, false, 1, d_syn, st);
        if (!rc && nt % K == 0 && nt / K < A->nck) rc = copy_state(c, A->ck[nt / K], S, st);
    }
    if (!rc) rc = probe_nan(c, S, N, st);
    if (!rc) rc = launch_ca_residual(d_syn, s->d_obs, A->res, s->ntr, N, A->J, st);
    for (int n = 0; n < 9 && !rc; ++n) {
        rc = ifos_fill(A->L[n], c->fe, 0.0f, st);
        if (!rc) rc = ifos_fill(A->mu_s[n], ifos_psi_elems(g, n), 0.0f, st);
        if (!rc) rc = ifos_fill(A->mu_v[n], ifos_psi_elems(g, n), 0.0f, st);
    }
    for (int n = 0; n < 8 && !rc; ++n) rc = ifos_fill((float *)A->G[n], 2 * c->fe, 0.0f, st);
    const ifos_model mod = ctx_model(c);
    for (int q = A->nck - 1; q >= 0 && !rc; --q) {
        const int a = q * K, b = std::min(a + K, N);
        rc = copy_state(c, A->work, A->ck[q], st);
        for (int n = a + 1; n <= b && !rc; ++n) {
            const size_t e = (size_t)(n - a - 1);
            for (int m = 0; m < 6 && !rc; ++m)
                rc = hip_check(hipMemcpyAsync(A->rs[6 * e + m], A->work.fields[3 + m], c->fe * 4,
                                              hipMemcpyDeviceToDevice, st), "ring copy");
            if (!rc) rc = step(c, A->work, 0, s, n, false, 1, nullptr, st);
            for (int m = 0; m < 3 && !rc; ++m)
                rc = hip_check(hipMemcpyAsync(A->rv[3 * e + m], A->work.fields[m], c->fe * 4, hipMemcpyDeviceToDevice,
                                              st), "ring copy");
            for (int m = 0; m < 9 && !rc; ++m) {
                const size_t pb = ifos_psi_elems(g, m) * 4;
                rc = hip_check(hipMemcpyAsync(A->rps[9 * e + m], A->work.psi_s[m], pb, hipMemcpyDeviceToDevice, st),
                               "ring copy");
                if (!rc)
                    rc = hip_check(hipMemcpyAsync(A->rpv[9 * e + m], A->work.psi_v[m], pb, hipMemcpyDeviceToDevice,
                                                  st), "ring copy");
            }
        }
        for (int n = b; n > a && !rc; --n) {
            const size_t e = (size_t)(n - a - 1);
            CaStep C;
            for (int m = 0; m < 9; ++m) {
                C.L[m] = A->L[m];
                C.scr[m] = A->scr[m];
                C.pv[m] = A->rpv[9 * e + m];
                C.ps[m] = A->rps[9 * e + m];
                C.mu_v[m] = A->mu_v[m];
                C.mu_s[m] = A->mu_s[m];
            }
            for (int m = 0; m < 3; ++m) C.v[m] = A->rv[3 * e + m];
            for (int m = 0; m < 6; ++m) C.s[m] = A->rs[6 * e + m];
            const float *mp[9] = {mod.rho, mod.pi, mod.u, mod.uipjp, mod.ujpkp, mod.uipkp, mod.rjp, mod.rkp, mod.rip};
            for (int m = 0; m < 9; ++m) C.model[m] = mp[m];
            for (int m = 0; m < 8; ++m) C.G[m] = A->G[m];
            C.rec = c->rec;
            C.ntr = s->ntr;
            C.res = A->res;
            C.n = n;
            rc = launch_ca_step(g, c->prof, C, st);
        }
    }
    if (!rc) rc = launch_ca_finish(g, A->G, &mod, A->Gr, A->Gp, A->Gu, d_grad, st);
    double hJ = 0.0;
    if (!rc) rc = hip_check(hipMemcpyAsync(&hJ, A->J, 8, hipMemcpyDeviceToHost, st), "misfit copy");
    if (!rc) rc = hip_check(hipStreamSynchronize(st), "consistent adjoint");
    if (!rc && J) *J = hJ;
    return rc;
}

// ---------------------------------------------------------------------------
// launch timer: each timed launch is bracketed by two events on its own
// stream.  Events come from a free list and go back to it once a pair has
// been read, so a long run (thousands of launches per propagation) holds at
// most 2 * (kMaxClosed + open scopes) events: when kMaxClosed pairs are
// waiting to be read, the oldest are waited for and folded into the sums.
namespace ifos {
struct ProfMark {
    int cat;
    hipEvent_t a, b;
};
struct ProfState {
    bool on = false;
    long next_id = 0;
    std::map<long, ProfMark> open;     // begun, end not yet recorded
    std::vector<ProfMark> closed;      // both events recorded, not yet read
    std::vector<hipEvent_t> free_ev;   // events ready for reuse
    double ms[IFOS_K_NCAT] = {};
    long count[IFOS_K_NCAT] = {};
};
static ProfState g_prof;
static const size_t kMaxClosed = 1024;

static hipEvent_t prof_event()
{
    if (!g_prof.free_ev.empty()) {
        hipEvent_t e = g_prof.free_ev.back();
        g_prof.free_ev.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// launches captured into a graph are not timed one by one
static bool capturing(hipStream_t s)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return s && hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

// reads every closed pair (waiting for its end event) and recycles its events
static void prof_fold_closed()
{
    for (auto &m : g_prof.closed) {
        float ms = 0.0f;
        if (hipEventSynchronize(m.b) == hipSuccess && hipEventElapsedTime(&ms, m.a, m.b) == hipSuccess) {
            g_prof.ms[m.cat] += ms;
            g_prof.count[m.cat] += 1;
        }
        g_prof.free_ev.push_back(m.a);
        g_prof.free_ev.push_back(m.b);
    }
    g_prof.closed.clear();
}

long prof_begin(int cat, hipStream_t s)
{
    if (!g_prof.on || capturing(s)) return -1;
    hipEvent_t a = prof_event();
    if (!a) return -1;
    hipEvent_t b = prof_event();
    if (!b) {
        g_prof.free_ev.push_back(a);
        return -1;
    }
    (void)hipEventRecord(a, s);
    const long id = g_prof.next_id++;
    g_prof.open[id] = ProfMark{cat, a, b};
    return id;
}

void prof_end(long id, hipStream_t s)
{
    if (id < 0) return;
    auto it = g_prof.open.find(id);
    if (it == g_prof.open.end()) return;  // dropped by a reset in between
    ProfMark m = it->second;
    g_prof.open.erase(it);
    (void)hipEventRecord(m.b, s);
    g_prof.closed.push_back(m);
    if (g_prof.closed.size() >= kMaxClosed) prof_fold_closed();
}

static void prof_drop_all()
{
    for (auto &kv : g_prof.open) {
        g_prof.free_ev.push_back(kv.second.a);
        g_prof.free_ev.push_back(kv.second.b);
    }
    g_prof.open.clear();
    for (auto &m : g_prof.closed) {
        g_prof.free_ev.push_back(m.a);
        g_prof.free_ev.push_back(m.b);
    }
    g_prof.closed.clear();
}
}  // namespace ifos

extern "C" int ifos_prof_enable(int on)
{
    g_prof.on = on != 0;
    return IFOS_OK;
}

extern "C" int ifos_prof_reset(void)
{
    (void)hipDeviceSynchronize();
    prof_drop_all();
    for (int n = 0; n < IFOS_K_NCAT; ++n) {
        g_prof.ms[n] = 0.0;
        g_prof.count[n] = 0;
    }
    return IFOS_OK;
}

extern "C" int ifos_prof_read(double *ms, long *count, int ncat)
{
    int rc = hip_check(hipDeviceSynchronize(), "ifos_prof_read sync");
    if (rc) return rc;
    prof_fold_closed();
    for (int n = 0; n < ncat && n < IFOS_K_NCAT; ++n) {
        if (ms) ms[n] = g_prof.ms[n];
        if (count) count[n] = g_prof.count[n];
    }
    return IFOS_OK;
}
