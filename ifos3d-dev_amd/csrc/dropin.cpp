// dropin.cpp -- libifos3d_dropin.so: the IFOS3D hot-path routines with the
// reference's own prototypes (include/ifos3d_dropin.h), running on the GPU
// through libifos3d_hip.so.
//
// The caller's NR arrays (util.c:250-400 layout: one block, k fastest) are
// mirrored on the device the first time a routine receives them, keyed by
// the float*** pointer, in the library's padded device layout.  The kind of
// an array (and so its index extents, ifos3d.c:288-348) follows from the
// parameter it is passed in.  Host copies are refreshed exactly where the
// reference driver reads them (see the header).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <unordered_map>
#include <vector>

#include "ifos3d_dropin.h"
#include "ifos3d_hip.h"

// the caller's globals (IFOS3D globvar.h:27-62); defined by the host program
extern "C" {
extern float DT, DX, DY, DZ, TIME, ALPHA, BETA, VPPML, FPML;
extern int NT, NX, NY, NZ, FDORDER, FDCOEFF, ABS_TYPE, FW, FREE_SURF, L, SEISMO, NDT, NDTSHIFT;
extern int POS[4], NPROCX, NPROCY, NPROCZ;
// the optimiser's (globvar.h:60-62)
extern int LBFGS, NUMPAR, BFGSNUM;
extern float VP0, VS0, RHO0, WEIGHT[3];
}

namespace {

enum Kind { WAVE_TOP, WAVE, MODEL, AVG, PSI_X, PSI_Y, PSI_Z, FREQ };

struct Mirror {
    Kind kind;
    float *dev = nullptr;
    int nf = 0;  // FREQ: frequencies held
};

struct State {
    bool have_geom = false;
    ifos_geom g;
    size_t fe = 0;
    std::unordered_map<const void *, Mirror> mir;
    // CPML profiles (18 vectors) on the device, and the host vector they came from
    float *d_prof = nullptr;
    const float *prof_key = nullptr;
    bool prof_dirty = true;
    // sources / receivers of the current propagation
    int *d_src = nullptr, *d_rec = nullptr, *d_chain = nullptr, *d_seisrec = nullptr;
    float *d_sig = nullptr, *d_adj = nullptr, *d_sec = nullptr;
    size_t cap_src = 0, cap_rec = 0, cap_sig = 0, cap_adj = 0, cap_seis = 0;
    const void *src_key[4] = {};  // srcpos, signals, signaly, signalz last uploaded
    int src_n = -1, src_back = -1;
    const void *seis_key = nullptr;
    int seis_ntr = -1;
    // samples gathered on the device and not yet copied to the host sections
    // (seismo), the sections they belong to and the layout [3][ntr][ns+1]
    int seis_ns = 0, seis_pending = 0;
    float **seis_out[3] = {};
    // the NaN probe's pinned slots (refresh_probe)
    float *probe_host = nullptr;
    hipEvent_t probe_ev[2] = {};
    bool probe_have[2] = {};
    // the L-BFGS history (bfgsgrad and bfgsscale live here; bfgsmod stays the
    // host's), keyed by the caller's bfgsgrad matrix
    ifos_lbfgs *lb = nullptr;
    const void *lb_key = nullptr;
    float *d_step = nullptr;  // modelupdate's bfgsmod slot
    size_t cap_step = 0;
    std::string err;
} S;

[[noreturn]] void fatal(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    S.err = buf;
    fprintf(stdout, "R U N - T I M E  E R R O R (ifos3d_dropin): %s\n...now exiting to system.\n", buf);
    fflush(stdout);
    exit(1);
}

void ck(int rc, const char *what)
{
    if (rc) fatal("%s: %s", what, ifos_last_error());
}

void hck(hipError_t e, const char *what)
{
    if (e != hipSuccess) fatal("%s: %s", what, hipGetErrorString(e));
}

// geometry from the caller's globals; the supported configuration is checked
const ifos_geom &geom()
{
    if (NPROCX * NPROCY * NPROCZ != 1 || POS[1] || POS[2] || POS[3])
        fatal("one MPI rank only (NPROCX*NPROCY*NPROCZ = %d)", NPROCX * NPROCY * NPROCZ);
    if (FDORDER != 4) fatal("FDORDER %d not supported (4 only)", FDORDER);
    if (L != 0 && L != 1) fatal("L = %d relaxation mechanisms not supported (0 or 1)", L);
    const ifos_geom &g = S.g;
    if (S.have_geom && g.nx == NX && g.ny == NY && g.nz == NZ && g.fw == FW && g.dx == DX && g.dy == DY &&
        g.dz == DZ && g.dt == DT && g.nt == NT && g.fdcoeff == FDCOEFF && g.free_surf == FREE_SURF)
        return g;
    if (S.have_geom && !S.mir.empty()) fatal("the grid changed while device mirrors exist");
    ck(ifos_geom_init(&S.g, NX, NY, NZ, FW, DX, DY, DZ, DT, NT, FDCOEFF, FREE_SURF), "grid");
    S.g.time = TIME;  // the caller's TIME (ifos3d.c:177 sets (NT-1)*DT)
    S.fe = ifos_field_elems(&S.g);
    S.have_geom = true;
    return S.g;
}

void extents(Kind k, int &jl, int &jh, int &il, int &ih, int &kl, int &kh)
{
    switch (k) {
    case WAVE_TOP: jl = -2; jh = NY + 2; il = -1; ih = NX + 2; kl = -1; kh = NZ + 2; break;
    case WAVE: jl = -1; jh = NY + 2; il = -1; ih = NX + 2; kl = -1; kh = NZ + 2; break;
    case MODEL: jl = 0; jh = NY + 1; il = 0; ih = NX + 1; kl = 0; kh = NZ + 1; break;
    default: jl = 1; jh = NY; il = 1; ih = NX; kl = 1; kh = NZ; break;
    }
}

size_t psi_elems(Kind k) { return ifos_psi_elems(&S.g, k == PSI_X ? 0 : k == PSI_Y ? 3 : 6); }

// host NR 3-D array <-> one device-layout field
void up3(float ***t, Kind k, float *dev)
{
    int jl, jh, il, ih, kl, kh;
    extents(k, jl, jh, il, ih, kl, kh);
    std::vector<float> st(S.fe, 0.0f);
    for (int j = jl; j <= jh; ++j)
        for (int i = il; i <= ih; ++i)
            memcpy(&st[ifos_offset(&S.g, j, i, kl)], &t[j][i][kl], sizeof(float) * (kh - kl + 1));
    hck(hipMemcpy(dev, st.data(), S.fe * 4, hipMemcpyHostToDevice), "upload");
}

void down3(float ***t, Kind k, const float *dev)
{
    int jl, jh, il, ih, kl, kh;
    extents(k, jl, jh, il, ih, kl, kh);
    std::vector<float> st(S.fe);
    hck(hipMemcpy(st.data(), dev, S.fe * 4, hipMemcpyDeviceToHost), "download");
    for (int j = jl; j <= jh; ++j)
        for (int i = il; i <= ih; ++i)
            memcpy(&t[j][i][kl], &st[ifos_offset(&S.g, j, i, kl)], sizeof(float) * (kh - kl + 1));
}

float *alloc_dev(size_t bytes)
{
    void *p = nullptr;
    hck(hipMalloc(&p, bytes ? bytes : 4), "hipMalloc");
    hck(hipMemset(p, 0, bytes ? bytes : 4), "hipMemset");
    return (float *)p;
}

// device mirror of a 3-D NR array (created, and filled from the host, on first sight)
float *m3(float ***t, Kind k, bool upload_new = true)
{
    auto it = S.mir.find(t);
    if (it != S.mir.end()) {
        if (it->second.kind != k) fatal("an array is passed as two different kinds");
        return it->second.dev;
    }
    Mirror m;
    m.kind = k;
    const bool psi = (k == PSI_X || k == PSI_Y || k == PSI_Z);
    const size_t elems = psi ? psi_elems(k) : S.fe;
    m.dev = alloc_dev(elems * 4);
    if (upload_new) {
        if (psi)  // compact [1..a][1..b][1..c] block, same shape on the device
            hck(hipMemcpy(m.dev, &t[1][1][1], elems * 4, hipMemcpyHostToDevice), "upload psi");
        else
            up3(t, k, m.dev);
    }
    S.mir[t] = m;
    return m.dev;
}

// device mirror of an f4tensor of DFT fields, frequencies 1..nf
float *m4(float ****t, int nf, bool upload_new = true)
{
    auto it = S.mir.find(t);
    if (it != S.mir.end() && it->second.nf >= nf) return it->second.dev;
    float *dev = alloc_dev(S.fe * 4 * (size_t)nf);
    int have = 0;
    if (it != S.mir.end()) {  // grow, keeping what the device holds
        have = it->second.nf;
        hck(hipMemcpy(dev, it->second.dev, S.fe * 4 * (size_t)have, hipMemcpyDeviceToDevice), "grow");
        hck(hipFree(it->second.dev), "hipFree");
    }
    if (upload_new)
        for (int m = have + 1; m <= nf; ++m) up3(t[m], WAVE_TOP, dev + (size_t)(m - 1) * S.fe);
    Mirror mm;
    mm.kind = FREQ;
    mm.dev = dev;
    mm.nf = nf;
    S.mir[t] = mm;
    return dev;
}

void grow(void **p, size_t &cap, size_t bytes)
{
    if (bytes <= cap) return;
    if (*p) hck(hipFree(*p), "hipFree");
    hck(hipMalloc(p, bytes), "hipMalloc");
    cap = bytes;
}

// sources of a propagation: forces (back 0) or adjoint traces at receivers
void upload_sources(float **srcpos, float **sx, float **sy, float **sz, int n, int back)
{
    const int nt = S.g.nt;
    if (back == 0) {
        std::vector<int> src(4 * (size_t)n);
        std::vector<float> sig((size_t)n * (nt + 2), 0.0f);
        for (int l = 1; l <= n; ++l) {
            src[4 * (l - 1) + 0] = (int)srcpos[1][l];
            src[4 * (l - 1) + 1] = (int)srcpos[2][l];
            src[4 * (l - 1) + 2] = (int)srcpos[3][l];
            src[4 * (l - 1) + 3] = (int)srcpos[7][l];
            memcpy(&sig[(size_t)(l - 1) * (nt + 2) + 1], &sx[l][1], sizeof(float) * nt);
        }
        grow((void **)&S.d_src, S.cap_src, src.size() * 4);
        grow((void **)&S.d_sig, S.cap_sig, sig.size() * 4);
        hck(hipMemcpy(S.d_src, src.data(), src.size() * 4, hipMemcpyHostToDevice), "upload sources");
        hck(hipMemcpy(S.d_sig, sig.data(), sig.size() * 4, hipMemcpyHostToDevice), "upload signals");
    } else {
        std::vector<int> rec(3 * (size_t)n), chain(2 * (size_t)n);
        std::vector<float> adj(3 * (size_t)n * (nt + 2), 0.0f);
        float **comp[3] = {sx, sy, sz};
        for (int l = 1; l <= n; ++l) {
            rec[3 * (l - 1) + 0] = (int)srcpos[1][l];
            rec[3 * (l - 1) + 1] = (int)srcpos[2][l];
            rec[3 * (l - 1) + 2] = (int)srcpos[3][l];
            for (int c = 0; c < 3; ++c)
                memcpy(&adj[((size_t)c * n + (l - 1)) * (nt + 2) + 1], &comp[c][l][1], sizeof(float) * nt);
        }
        ck(ifos_rec_chain(rec.data(), n, chain.data()), "receiver chain");
        grow((void **)&S.d_rec, S.cap_rec, rec.size() * 4 + chain.size() * 4);
        S.d_chain = S.d_rec + rec.size();
        grow((void **)&S.d_adj, S.cap_adj, adj.size() * 4);
        hck(hipMemcpy(S.d_rec, rec.data(), rec.size() * 4, hipMemcpyHostToDevice), "upload receivers");
        hck(hipMemcpy(S.d_chain, chain.data(), chain.size() * 4, hipMemcpyHostToDevice), "upload chain");
        hck(hipMemcpy(S.d_adj, adj.data(), adj.size() * 4, hipMemcpyHostToDevice), "upload traces");
    }
    S.src_key[0] = srcpos;
    S.src_key[1] = sx;
    S.src_key[2] = sy;
    S.src_key[3] = sz;
    S.src_n = n;
    S.src_back = back;
}

// force: the caller may have rewritten the arrays (first step of a propagation)
void ensure_sources(bool force, float **srcpos, float **sx, float **sy, float **sz, int n, int back)
{
    if (n <= 0) return;
    if (force || S.src_back != back || S.src_n != n || S.src_key[0] != srcpos || S.src_key[1] != sx ||
        S.src_key[2] != sy || S.src_key[3] != sz)
        upload_sources(srcpos, sx, sy, sz, n, back);
}

// the 18 kernel-visible CPML vectors (1-based [1..2FW]) into the device array
void ensure_profiles(float *const v[18])
{
    if (!S.prof_dirty && S.prof_key == v[0] && S.d_prof) return;
    const int len = 2 * FW + 1;
    std::vector<float> h(18 * (size_t)len, 0.0f);
    for (int n = 0; n < 18; ++n) memcpy(&h[(size_t)n * len + 1], &v[n][1], sizeof(float) * 2 * FW);
    if (!S.d_prof) S.d_prof = alloc_dev(h.size() * 4);
    hck(hipMemcpy(S.d_prof, h.data(), h.size() * 4, hipMemcpyHostToDevice), "upload CPML profiles");
    S.prof_key = v[0];
    S.prof_dirty = false;
}

void check_box(int nx1, int nx2, int ny1, int ny2, int nz1, int nz2)
{
    int xb[2], yb[2], zb[2];
    CPML_ini_elastic(xb, yb, zb);
    if (nx1 != xb[0] || nx2 != xb[1] || ny1 != yb[0] || ny2 != yb[1] || nz1 != zb[0] || nz2 != zb[1])
        fatal("CPML slabs need the CPML_ini_elastic box, got [%d,%d]x[%d,%d]x[%d,%d]", nx1, nx2, ny1, ny2, nz1, nz2);
}

// The host NaN probe of ifos3d.c:761 reads vy[NY/2][NX/2][NZ/2] before each
// step.  After update_v_CPML of step nt (the step's last velocity update) the
// cell is copied, asynchronously, into one of two pinned host slots; at step
// nt + 1 the slot of step nt is written into the caller's array.  No call
// waits for the device (a wait per step would leave the device idle while the
// host issues the next launches); the host's check at the top of step nt + 2
// sees step nt's value, so an unstable run stops one step later than the
// reference's.  Step 1 reads no slot (the previous propagation's).
void refresh_probe(float ***vy, const float *dvy, int nt)
{
    if (!S.probe_host) {
        hck(hipHostMalloc((void **)&S.probe_host, 2 * sizeof(float), hipHostMallocDefault), "probe slots");
        for (int q = 0; q < 2; ++q)
            hck(hipEventCreateWithFlags(&S.probe_ev[q], hipEventDisableTiming), "probe events");
    }
    const int cur = nt & 1, prev = cur ^ 1;
    if (nt > 1 && S.probe_have[prev]) {
        hck(hipEventSynchronize(S.probe_ev[prev]), "probe wait");
        vy[NY / 2][NX / 2][NZ / 2] = S.probe_host[prev];
    }
    S.probe_have[prev] = false;
    const size_t o = ifos_offset(&S.g, NY / 2, NX / 2, NZ / 2);
    hck(hipMemcpyAsync(&S.probe_host[cur], dvy + o, 4, hipMemcpyDeviceToHost, nullptr), "NaN probe");
    hck(hipEventRecord(S.probe_ev[cur], nullptr), "probe event");
    S.probe_have[cur] = true;
}

ifos_fields fields(float ***vx, float ***vy, float ***vz, float ***sxx, float ***syy, float ***szz, float ***sxy,
                   float ***syz, float ***sxz)
{
    ifos_fields f;
    f.vx = m3(vx, WAVE_TOP);
    f.vy = m3(vy, WAVE_TOP);
    f.vz = m3(vz, WAVE_TOP);
    f.sxx = m3(sxx, WAVE);
    f.syy = m3(syy, WAVE);
    f.szz = m3(szz, WAVE);
    f.sxy = m3(sxy, WAVE_TOP);
    f.syz = m3(syz, WAVE_TOP);
    f.sxz = m3(sxz, WAVE);
    return f;
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

double update_v(int nx1, int nx2, int ny1, int ny2, int nz1, int nz2, int nt, float ***vx, float ***vy, float ***vz,
                float ***sxx, float ***syy, float ***szz, float ***sxy, float ***syz, float ***sxz, float ***rho,
                float ***rjp, float ***rkp, float ***rip, float **srcpos_loc, float **signals, float **signaly,
                float **signalz, int nsrc, float ***absorb_coeff, int back)
{
    (void)rho;
    (void)absorb_coeff;
    const ifos_geom &g = geom();
    if (ABS_TYPE != 1) fatal("ABS_TYPE %d not supported (CPML only)", ABS_TYPE);
    ifos_fields f = fields(vx, vy, vz, sxx, syy, szz, sxy, syz, sxz);
    ifos_model m = {};
    m.rjp = m3(rjp, AVG);
    m.rkp = m3(rkp, AVG);
    m.rip = m3(rip, AVG);
    const int box[6] = {nx1, nx2, ny1, ny2, nz1, nz2};
    ck(ifos_update_v_box(&g, box, &f, &m, nullptr), "update_v");
    ensure_sources(nt == 1, srcpos_loc, signals, signaly, signalz, nsrc, back);
    if (nsrc > 0) {
        if (back == 0)
            ck(ifos_add_forces(&g, &f, &m, S.d_src, S.d_sig, nsrc, nt, ALPHA, BETA, nullptr), "forces");
        else
            ck(ifos_add_adjoint(&g, &f, S.d_rec, S.d_chain, S.d_adj, nsrc, nt, nullptr), "adjoint traces");
    }
    return 0.0;
}

double update_v_CPML(int nx1, int nx2, int ny1, int ny2, int nz1, int nz2, int nt, float ***vx, float ***vy,
                     float ***vz, float ***sxx, float ***syy, float ***szz, float ***sxy, float ***syz, float ***sxz,
                     float ***rho, float ***rjp, float ***rkp, float ***rip, float **srcpos_loc, float **signals,
                     float **signaly, float **signalz, int nsrc, float ***absorb_coeff, int back, float *K_x,
                     float *a_x, float *b_x, float *K_x_half, float *a_x_half, float *b_x_half, float *K_y,
                     float *a_y, float *b_y, float *K_y_half, float *a_y_half, float *b_y_half, float *K_z,
                     float *a_z, float *b_z, float *K_z_half, float *a_z_half, float *b_z_half, float ***psi_sxx_x,
                     float ***psi_sxy_x, float ***psi_sxz_x, float ***psi_sxy_y, float ***psi_syy_y,
                     float ***psi_syz_y, float ***psi_sxz_z, float ***psi_syz_z, float ***psi_szz_z)
{
    (void)rho;
    (void)absorb_coeff;
    const ifos_geom &g = geom();
    check_box(nx1, nx2, ny1, ny2, nz1, nz2);
    ifos_fields f = fields(vx, vy, vz, sxx, syy, szz, sxy, syz, sxz);
    ifos_model m = {};
    m.rjp = m3(rjp, AVG);
    m.rkp = m3(rkp, AVG);
    m.rip = m3(rip, AVG);
    float *const prof[18] = {K_x, a_x, b_x, K_x_half, a_x_half, b_x_half, K_y, a_y, b_y,
                             K_y_half, a_y_half, b_y_half, K_z, a_z, b_z, K_z_half, a_z_half, b_z_half};
    if (nt == 1) S.prof_dirty = true;
    ensure_profiles(prof);
    ifos_psi psi;
    float ***p[9] = {psi_sxx_x, psi_sxy_x, psi_sxz_x, psi_sxy_y, psi_syy_y, psi_syz_y, psi_sxz_z, psi_syz_z, psi_szz_z};
    for (int n = 0; n < 9; ++n) psi.p[n] = m3(p[n], n < 3 ? PSI_X : n < 6 ? PSI_Y : PSI_Z);
    ck(ifos_update_v(&g, IFOS_REGION_CPML, &f, &m, S.d_prof, &psi, nullptr), "update_v_CPML");
    if (back == 1 && nsrc > 0) {  // update_v_ssg_CPML.c:384-393 injects the traces again
        ensure_sources(false, srcpos_loc, signals, signaly, signalz, nsrc, back);
        ck(ifos_add_adjoint(&g, &f, S.d_rec, S.d_chain, S.d_adj, nsrc, nt, nullptr), "adjoint traces");
    }
    refresh_probe(vy, f.vy, nt);
    return 0.0;
}

double update_s_elastic(int nx1, int nx2, int ny1, int ny2, int nz1, int nz2, float ***vx, float ***vy, float ***vz,
                        float ***sxx, float ***syy, float ***szz, float ***sxy, float ***syz, float ***sxz, float ***pi,
                        float ***u, float ***uipjp, float ***ujpkp, float ***uipkp)
{
    const ifos_geom &g = geom();
    ifos_fields f = fields(vx, vy, vz, sxx, syy, szz, sxy, syz, sxz);
    ifos_model m = {};
    m.pi = m3(pi, MODEL);
    m.u = m3(u, MODEL);
    m.uipjp = m3(uipjp, AVG);
    m.ujpkp = m3(ujpkp, AVG);
    m.uipkp = m3(uipkp, AVG);
    const int box[6] = {nx1, nx2, ny1, ny2, nz1, nz2};
    ck(ifos_update_s_box(&g, box, &f, &m, nullptr), "update_s_elastic");
    return 0.0;
}

double update_s_CPML_elastic(int nx1, int nx2, int ny1, int ny2, int nz1, int nz2, float ***vx, float ***vy,
                             float ***vz, float ***sxx, float ***syy, float ***szz, float ***sxy, float ***syz,
                             float ***sxz, float ***pi, float ***u, float ***uipjp, float ***ujpkp, float ***uipkp,
                             float *K_x, float *a_x, float *b_x, float *K_x_half, float *a_x_half, float *b_x_half,
                             float *K_y, float *a_y, float *b_y, float *K_y_half, float *a_y_half, float *b_y_half,
                             float *K_z, float *a_z, float *b_z, float *K_z_half, float *a_z_half, float *b_z_half,
                             float ***psi_vxx, float ***psi_vyx, float ***psi_vzx, float ***psi_vxy, float ***psi_vyy,
                             float ***psi_vzy, float ***psi_vxz, float ***psi_vyz, float ***psi_vzz)
{
    const ifos_geom &g = geom();
    check_box(nx1, nx2, ny1, ny2, nz1, nz2);
    ifos_fields f = fields(vx, vy, vz, sxx, syy, szz, sxy, syz, sxz);
    ifos_model m = {};
    m.pi = m3(pi, MODEL);
    m.u = m3(u, MODEL);
    m.uipjp = m3(uipjp, AVG);
    m.ujpkp = m3(ujpkp, AVG);
    m.uipkp = m3(uipkp, AVG);
    float *const prof[18] = {K_x, a_x, b_x, K_x_half, a_x_half, b_x_half, K_y, a_y, b_y,
                             K_y_half, a_y_half, b_y_half, K_z, a_z, b_z, K_z_half, a_z_half, b_z_half};
    ensure_profiles(prof);
    ifos_psi psi;
    float ***p[9] = {psi_vxx, psi_vyx, psi_vzx, psi_vxy, psi_vyy, psi_vzy, psi_vxz, psi_vyz, psi_vzz};
    for (int n = 0; n < 9; ++n) psi.p[n] = m3(p[n], n < 3 ? PSI_X : n < 6 ? PSI_Y : PSI_Z);
    ck(ifos_update_s(&g, IFOS_REGION_CPML, &f, &m, S.d_prof, &psi, nullptr), "update_s_CPML_elastic");
    return 0.0;
}

namespace {
ifos_visco visco_args(float ***rxx, float ***ryy, float ***rzz, float ***rxy, float ***ryz, float ***rxz,
                      float ***taus, float ***tausipjp, float ***tausjpkp, float ***tausipkp, float ***taup,
                      float *eta)
{
    if (L != 1) fatal("update_s / update_s_CPML need L = 1 (L = %d)", L);
    ifos_visco V = {};
    float ***r[6] = {rxx, ryy, rzz, rxy, ryz, rxz};
    for (int n = 0; n < 6; ++n) V.r[n] = m3(r[n], AVG);
    V.taus = m3(taus, MODEL);
    V.taup = m3(taup, MODEL);
    V.tausipjp = m3(tausipjp, AVG);
    V.tausjpkp = m3(tausjpkp, AVG);
    V.tausipkp = m3(tausipkp, AVG);
    V.eta = eta[1];
    V.L = L;
    return V;
}
}  // namespace

double update_s(int nx1, int nx2, int ny1, int ny2, int nz1, int nz2, float ***vx, float ***vy, float ***vz,
                float ***sxx, float ***syy, float ***szz, float ***sxy, float ***syz, float ***sxz, float ***rxx,
                float ***ryy, float ***rzz, float ***rxy, float ***ryz, float ***rxz, float ***pi, float ***u,
                float ***uipjp, float ***ujpkp, float ***uipkp, float ***taus, float ***tausipjp, float ***tausjpkp,
                float ***tausipkp, float ***taup, float *eta)
{
    const ifos_geom &g = geom();
    check_box(nx1, nx2, ny1, ny2, nz1, nz2);
    ifos_fields f = fields(vx, vy, vz, sxx, syy, szz, sxy, syz, sxz);
    ifos_model m = {};
    m.pi = m3(pi, MODEL);
    m.u = m3(u, MODEL);
    m.uipjp = m3(uipjp, AVG);
    m.ujpkp = m3(ujpkp, AVG);
    m.uipkp = m3(uipkp, AVG);
    const ifos_visco V = visco_args(rxx, ryy, rzz, rxy, ryz, rxz, taus, tausipjp, tausjpkp, tausipkp, taup, eta);
    ck(ifos_update_s_visco(&g, IFOS_REGION_INTERIOR, &f, &m, &V, nullptr, nullptr, nullptr), "update_s");
    return 0.0;
}

double update_s_CPML(int nx1, int nx2, int ny1, int ny2, int nz1, int nz2, float ***vx, float ***vy, float ***vz,
                     float ***sxx, float ***syy, float ***szz, float ***sxy, float ***syz, float ***sxz, float ***rxx,
                     float ***ryy, float ***rzz, float ***rxy, float ***ryz, float ***rxz, float ***pi, float ***u,
                     float ***uipjp, float ***ujpkp, float ***uipkp, float ***taus, float ***tausipjp,
                     float ***tausjpkp, float ***tausipkp, float ***taup, float *eta, float *K_x, float *a_x,
                     float *b_x, float *K_x_half, float *a_x_half, float *b_x_half, float *K_y, float *a_y, float *b_y,
                     float *K_y_half, float *a_y_half, float *b_y_half, float *K_z, float *a_z, float *b_z,
                     float *K_z_half, float *a_z_half, float *b_z_half, float ***psi_vxx, float ***psi_vyx,
                     float ***psi_vzx, float ***psi_vxy, float ***psi_vyy, float ***psi_vzy, float ***psi_vxz,
                     float ***psi_vyz, float ***psi_vzz)
{
    const ifos_geom &g = geom();
    check_box(nx1, nx2, ny1, ny2, nz1, nz2);
    ifos_fields f = fields(vx, vy, vz, sxx, syy, szz, sxy, syz, sxz);
    ifos_model m = {};
    m.pi = m3(pi, MODEL);
    m.u = m3(u, MODEL);
    m.uipjp = m3(uipjp, AVG);
    m.ujpkp = m3(ujpkp, AVG);
    m.uipkp = m3(uipkp, AVG);
    float *const prof[18] = {K_x, a_x, b_x, K_x_half, a_x_half, b_x_half, K_y, a_y, b_y,
                             K_y_half, a_y_half, b_y_half, K_z, a_z, b_z, K_z_half, a_z_half, b_z_half};
    ensure_profiles(prof);
    ifos_psi psi;
    float ***p[9] = {psi_vxx, psi_vyx, psi_vzx, psi_vxy, psi_vyy, psi_vzy, psi_vxz, psi_vyz, psi_vzz};
    for (int n = 0; n < 9; ++n) psi.p[n] = m3(p[n], n < 3 ? PSI_X : n < 6 ? PSI_Y : PSI_Z);
    const ifos_visco V = visco_args(rxx, ryy, rzz, rxy, ryz, rxz, taus, tausipjp, tausjpkp, tausipkp, taup, eta);
    ck(ifos_update_s_visco(&g, IFOS_REGION_CPML, &f, &m, &V, S.d_prof, &psi, nullptr), "update_s_CPML");
    return 0.0;
}

void psource(int nt, float ***sxx, float ***syy, float ***szz, float **srcpos_loc, float **signals, int nsrc)
{
    const ifos_geom &g = geom();
    if (nsrc <= 0) return;
    ensure_sources(false, srcpos_loc, signals, signals, signals, nsrc, 0);
    ifos_fields f = {};
    f.sxx = m3(sxx, WAVE);
    f.syy = m3(syy, WAVE);
    f.szz = m3(szz, WAVE);
    ck(ifos_psource(&g, &f, S.d_src, S.d_sig, nsrc, nt, nullptr), "psource");
}

namespace {
// copies samples 1..seis_pending of the device sections [3][ntr][ns+1] into
// the host sections they were gathered for
void flush_seis()
{
    if (S.seis_pending <= 0) return;
    const int ntr = S.seis_ntr, ld = S.seis_ns + 1;
    std::vector<float> h(3 * (size_t)ntr * ld);
    hck(hipMemcpy(h.data(), S.d_sec, h.size() * 4, hipMemcpyDeviceToHost), "download sections");
    for (int c = 0; c < 3; ++c)
        for (int r = 1; r <= ntr; ++r)
            memcpy(&S.seis_out[c][r][1], &h[((size_t)c * ntr + r - 1) * ld + 1], sizeof(float) * S.seis_pending);
    S.seis_pending = 0;
}
}  // namespace

// seismo_ssg.c:27-53 (SEISMO 1).  Sample lsamp of every receiver is gathered
// on the device; the samples of a propagation reach the host sections in one
// copy, at its last sample ns = ceil((NT - NDTSHIFT) / NDT) (ifos3d.c:178-179),
// before saveseis and the residual read them (ifos3d.c:816, 839-858) -- and
// before another propagation (lsamp 1) or other receivers start.
void seismo(int lsamp, int ntr, int **recpos, float **sectionvx, float **sectionvy, float **sectionvz,
            float **sectiondiv, float **sectioncurl, float **sectionp, float ***vx, float ***vy, float ***vz,
            float ***sxx, float ***syy, float ***szz, float ***pi, float ***u)
{
    (void)sectiondiv, (void)sectioncurl, (void)sectionp, (void)sxx, (void)syy, (void)szz, (void)pi, (void)u;
    const ifos_geom &g = geom();
    if (SEISMO != 1) fatal("SEISMO %d not supported (particle velocities, 1, only)", SEISMO);
    if (ntr <= 0) return;
    const int ns = NDTSHIFT > NT || NDT < 1 ? 0 : (int)ceil((float)(NT - NDTSHIFT) / (float)NDT);
    float **out[3] = {sectionvx, sectionvy, sectionvz};
    if (lsamp < 1 || lsamp > ns) fatal("seismo: sample %d outside 1..%d (NT %d, NDT %d, NDTSHIFT %d)", lsamp, ns, NT,
                                       NDT, NDTSHIFT);
    const bool same = S.seis_key == recpos && S.seis_ntr == ntr && S.seis_ns == ns &&
                      S.seis_out[0] == sectionvx && S.seis_out[1] == sectionvy && S.seis_out[2] == sectionvz;
    if (lsamp == 1 || !same) {
        flush_seis();
        std::vector<int> rec(3 * (size_t)ntr);
        for (int r = 1; r <= ntr; ++r) {
            rec[3 * (r - 1) + 0] = recpos[1][r];
            rec[3 * (r - 1) + 1] = recpos[2][r];
            rec[3 * (r - 1) + 2] = recpos[3][r];
        }
        grow((void **)&S.d_seisrec, S.cap_seis, rec.size() * 4 + 3 * (size_t)ntr * (ns + 1) * 4);
        S.d_sec = (float *)(S.d_seisrec + rec.size());
        hck(hipMemcpy(S.d_seisrec, rec.data(), rec.size() * 4, hipMemcpyHostToDevice), "upload receivers");
        S.seis_key = recpos;
        S.seis_ntr = ntr;
        S.seis_ns = ns;
        for (int c = 0; c < 3; ++c) S.seis_out[c] = out[c];
        S.seis_pending = 0;
    }
    ifos_fields f = {};
    f.vx = m3(vx, WAVE_TOP);
    f.vy = m3(vy, WAVE_TOP);
    f.vz = m3(vz, WAVE_TOP);
    ck(ifos_seismo(&g, &f, S.d_seisrec, ntr, lsamp, ns, S.d_sec, nullptr), "seismo");
    if (lsamp > S.seis_pending) S.seis_pending = lsamp;
    if (lsamp == ns) flush_seis();
}

void surface_elastic(int ndepth, float ***u, float ***pi, float ***sxx, float ***syy, float ***szz, float ***sxy,
                     float ***syz, float ***vx, float ***vy, float ***vz)
{
    const ifos_geom &g = geom();
    if (ndepth != 1) fatal("surface_elastic: free surface at j = %d not supported (1 only)", ndepth);
    ifos_fields f = {};
    f.vx = m3(vx, WAVE_TOP);
    f.vy = m3(vy, WAVE_TOP);
    f.vz = m3(vz, WAVE_TOP);
    f.sxx = m3(sxx, WAVE);
    f.syy = m3(syy, WAVE);
    f.szz = m3(szz, WAVE);
    f.sxy = m3(sxy, WAVE_TOP);
    f.syz = m3(syz, WAVE_TOP);
    ifos_model m = {};
    m.pi = m3(pi, MODEL);
    m.u = m3(u, MODEL);
    ck(ifos_surface(&g, &f, &m, nullptr), "surface_elastic");
}

void discfourier(int nx1, int nx2, int ny1, int ny2, int nz1, int nz2, int nt, float ***vx, float ***vy, float ***vz,
                 float ****Fvx, float ****Fvy, float ****Fvz, float ****Fvxi, float ****Fvyi, float ****Fvzi,
                 float *finv, int nf, int ntast, int pshot, int back)
{
    const ifos_geom &g = geom();
    if (nx1 != 1 || ny1 != 1 || nz1 != 1 || nx2 != NX || ny2 != NY || nz2 != NZ)
        fatal("discfourier over a sub-box not supported (whole local grid only)");
    if (pshot < 0) fatal("discfourier: pshot %d", pshot);
    ifos_fields f = {};
    f.vx = m3(vx, WAVE_TOP);
    f.vy = m3(vy, WAVE_TOP);
    f.vz = m3(vz, WAVE_TOP);
    float ****F[6] = {Fvx, Fvy, Fvz, Fvxi, Fvyi, Fvzi};
    float *d[6];
    // disc_fourier.c:40: frequency l of Hessian wavefield pshot is entry pshot*nf + l
    for (int n = 0; n < 6; ++n) d[n] = m4(F[n], (pshot + 1) * nf) + (size_t)pshot * nf * S.fe;
    ck(ifos_discfourier(&g, &f, d, nullptr, finv, nf, nt, ntast, back, IFOS_DFT_COMPAT, nullptr), "discfourier");
}

void gradient_F(int nx, int ny, int nz, float ****fvx, float ****fvy, float ****fvz, float ****fivx, float ****fivy,
                float ****fivz, float ****bvx, float ****bvy, float ****bvz, float ****bivx, float ****bivy,
                float ****bivz, float ***gradlam, float ***gradmu, float ***gradrho, int nt, float ***rho,
                float ***pi, float ***u, float *finv, int nf, int iteration)
{
    (void)nt;
    (void)iteration;
    const ifos_geom &g = geom();
    if (nx != NX || ny != NY || nz != NZ) fatal("gradient_F: grid %dx%dx%d != local grid", nx, ny, nz);
    float ****Ff[6] = {fvx, fvy, fvz, fivx, fivy, fivz}, ****Fb[6] = {bvx, bvy, bvz, bivx, bivy, bivz};
    float *df[6], *db[6];
    for (int n = 0; n < 6; ++n) {
        df[n] = m4(Ff[n], nf);
        db[n] = m4(Fb[n], nf);
    }
    ifos_model m = {};
    m.rho = m3(rho, MODEL);
    m.pi = m3(pi, MODEL);
    m.u = m3(u, MODEL);
    float ***G[3] = {gradlam, gradmu, gradrho};
    float *dg[3];
    for (int n = 0; n < 3; ++n) {  // the host may have changed the gradients (zero_grad, outgrad)
        dg[n] = m3(G[n], AVG, false);
        up3(G[n], AVG, dg[n]);
    }
    ck(ifos_gradient_F(&g, df, db, &m, finv, nf, dg, nullptr), "gradient_F");
    for (int n = 0; n < 3; ++n) down3(G[n], AVG, dg[n]);
}

void hess_F(int nx, int ny, int nz, float ****fvx, float ****fvy, float ****fvz, float ****fivx, float ****fivy,
            float ****fivz, float ****bvx, float ****bvy, float ****bvz, float ****bivx, float ****bivy,
            float ****bivz, float ***hess1, float ***hess2, float ***hess3, int nt, float ***rho, float ***pi,
            float ***u, float *finv, int nf, int ntr_hess)
{
    (void)nt;
    const ifos_geom &g = geom();
    if (nx != NX || ny != NY || nz != NZ) fatal("hess_F: grid %dx%dx%d != local grid", nx, ny, nz);
    if (ntr_hess < 0) fatal("hess_F: ntr_hess %d", ntr_hess);
    float ****Ff[6] = {fvx, fvy, fvz, fivx, fivy, fivz}, ****Fb[6] = {bvx, bvy, bvz, bivx, bivy, bivz};
    float *df[6], *db[6];
    for (int n = 0; n < 6; ++n) {
        df[n] = m4(Ff[n], nf);
        db[n] = m4(Fb[n], nf * (ntr_hess + 1));
    }
    ifos_model m = {};
    m.rho = m3(rho, MODEL);
    m.pi = m3(pi, MODEL);
    m.u = m3(u, MODEL);
    float ***H[3] = {hess1, hess2, hess3};
    float *dh[3];
    for (int n = 0; n < 3; ++n) {  // the host may have changed them (readhess, outgrad)
        dh[n] = m3(H[n], AVG, false);
        up3(H[n], AVG, dh[n]);
    }
    ck(ifos_hess_F(&g, df, db, ntr_hess, &m, finv, nf, dh, nullptr), "hess_F");
    for (int n = 0; n < 3; ++n) down3(H[n], AVG, dh[n]);
}

// ---- optimiser steps (lbfgs_save.c, lbfgs.c, modelupdate.c, conjugategrad.c) on the device:
// the gradients, priors and model are uploaded from the host arrays at each
// call and written back (the host reads them in between: outgrad, matcopy,
// outmod); bfgsgrad and bfgsscale are kept on the device in FP64 sums,
// bfgsmod is the host's (modelupdate writes it there, lbfgs uploads it)
static ifos_lbfgs *lb_for(float **bfgsgrad)
{
    const ifos_geom &g = geom();
    if (!S.lb || S.lb_key != bfgsgrad) {
        if (S.lb) ifos_lbfgs_destroy(S.lb);
        S.lb = nullptr;
        if (NUMPAR < 1 || NUMPAR > 2 || BFGSNUM < 1) fatal("L-BFGS: NUMPAR %d, BFGSNUM %d", NUMPAR, BFGSNUM);
        ck(ifos_lbfgs_create(&g, NUMPAR, BFGSNUM, &S.lb), "lbfgs");
        S.lb_key = bfgsgrad;
    }
    return S.lb;
}

static float *upload_avg(float ***t)
{
    float *d = m3(t, AVG, false);
    up3(t, AVG, d);
    return d;
}

void lbfgs_savegrad(float ***grad1, float ***grad2, float ***grad3, float **bfgsgrad1)
{
    (void)grad3;
    ifos_lbfgs *b = lb_for(bfgsgrad1);
    ck(ifos_lbfgs_savegrad(b, upload_avg(grad1), NUMPAR > 1 ? upload_avg(grad2) : nullptr), "lbfgs_savegrad");
}

void lbfgs(float ***grad1, float ***grad2, float ***grad3, float *bfgsscale, float **bfgsmod, float **bfgsgrad,
           int iteration)
{
    (void)grad3;
    ifos_lbfgs *b = lb_for(bfgsgrad);
    const size_t n = (size_t)NUMPAR * NX * NY * NZ;
    for (int w = 1; w <= BFGSNUM; ++w)  // the model steps modelupdate wrote
        hck(hipMemcpy(ifos_lbfgs_mod(b, w), &bfgsmod[w][1], n * 4, hipMemcpyHostToDevice), "upload bfgsmod");
    float *d1 = upload_avg(grad1), *d2 = NUMPAR > 1 ? upload_avg(grad2) : nullptr;
    ck(ifos_lbfgs_update(b, d1, d2, iteration, nullptr), "lbfgs");
    down3(grad1, AVG, d1);
    if (d2) down3(grad2, AVG, d2);
    std::vector<double> sc(BFGSNUM);
    ck(ifos_lbfgs_scale(b, sc.data()), "lbfgs");
    for (int w = 1; w <= BFGSNUM; ++w) bfgsscale[w] = (float)sc[w - 1];
}

void modelupdate(int nx, int ny, int nz, float ***gradvp, float ***gradvs, float ***gradrho, float ***rho,
                 float ***pi, float ***u, float **bfgsmod1, float step, float *beta, int it_group)
{
    (void)beta;
    const ifos_geom &g = geom();
    if (nx != NX || ny != NY || nz != NZ) fatal("modelupdate: grid %dx%dx%d != local grid", nx, ny, nz);
    float ***G[3] = {gradvp, gradvs, gradrho}, ***M[3] = {rho, pi, u};
    float *dg[3], *dm[3];
    for (int q = 0; q < 3; ++q) {
        dg[q] = upload_avg(G[q]);
        dm[q] = m3(M[q], MODEL, false);  // the caller's copy (cpmodel) or the model itself
        up3(M[q], MODEL, dm[q]);
    }
    const size_t n = (size_t)NUMPAR * NX * NY * NZ;
    if (LBFGS) grow((void **)&S.d_step, S.cap_step, n * 4);
    ck(ifos_modelupdate(&g, dg, dm[0], dm[1], dm[2], LBFGS ? S.d_step : nullptr, NUMPAR, LBFGS, step, WEIGHT, VP0,
                        VS0, RHO0, nullptr, nullptr),
       "modelupdate");
    for (int q = 0; q < 3; ++q) {
        down3(M[q], MODEL, dm[q]);
        down3(G[q], AVG, dg[q]);  // without L-BFGS the gradients were normalised in place
    }
    if (LBFGS) {
        int w = it_group % BFGSNUM;
        if (w == 0) w = BFGSNUM;
        hck(hipMemcpy(&bfgsmod1[w][1], S.d_step, n * 4, hipMemcpyDeviceToHost), "download bfgsmod");
    }
}

void conjugate(int nx, int ny, int nz, float ***grad1, float ***grad2, float ***grad3, float ***gradprior1,
               float ***gradprior2, float ***gradprior3, float ***gradprior4, float ***gradprior5,
               float ***gradprior6, float *beta, int iteration, int cdf)
{
    (void)iteration;
    const ifos_geom &g = geom();
    if (nx != NX || ny != NY || nz != NZ) fatal("conjugate: grid %dx%dx%d != local grid", nx, ny, nz);
    float ***G[3] = {grad1, grad2, grad3};
    float ***P[6] = {gradprior1, gradprior2, gradprior3, gradprior4, gradprior5, gradprior6};
    float *dg[3], *dp[6];
    for (int q = 0; q < 3; ++q) dg[q] = upload_avg(G[q]);
    for (int q = 0; q < 6; ++q) dp[q] = upload_avg(P[q]);
    float b3[3] = {0.0f, 0.0f, 0.0f};
    ck(ifos_conjugate(&g, dg, dp, b3, cdf, nullptr), "conjugate");
    for (int q = 0; q < 3; ++q) down3(G[q], AVG, dg[q]);
    for (int q = 0; q < 6; ++q) down3(P[q], AVG, dp[q]);
    for (int q = 0; q < 3; ++q) beta[q] = b3[q];
}

void av_mat(float ***rho, float ***pi, float ***u, float ***taus, float ***taup, float ***uipjp, float ***ujpkp,
            float ***uipkp, float ***tausipjp, float ***tausjpkp, float ***tausipkp, float ***rjp, float ***rkp,
            float ***rip)
{
    const ifos_geom &g = geom();
    // the model (and its halos from the caller's matcopy) is re-read every call
    float ***in[3] = {rho, pi, u};
    float *d[3];
    for (int n = 0; n < 3; ++n) {
        d[n] = m3(in[n], MODEL, false);
        up3(in[n], MODEL, d[n]);
    }
    float ***out[6] = {uipjp, ujpkp, uipkp, rjp, rkp, rip};
    float *o[6];
    for (int n = 0; n < 6; ++n) o[n] = m3(out[n], AVG, false);
    ck(ifos_av_mat(&g, d[0], d[2], o[0], o[1], o[2], o[3], o[4], o[5], nullptr), "av_mat");
    for (int n = 0; n < 6; ++n) down3(out[n], AVG, o[n]);
    if (L) {  // av_mat.c:56-60, and the relaxation ratios are (re)read like the model
        float *dts = m3(taus, MODEL, false), *dtp = m3(taup, MODEL, false);
        up3(taus, MODEL, dts);
        up3(taup, MODEL, dtp);
        float ***tout[3] = {tausipjp, tausjpkp, tausipkp};
        float *t[3];
        for (int n = 0; n < 3; ++n) t[n] = m3(tout[n], AVG, false);
        ck(ifos_av_tau(&g, dts, t[0], t[1], t[2], nullptr), "av_mat");
        for (int n = 0; n < 3; ++n) down3(tout[n], AVG, t[n]);
    }
}

void CPML_coeff(float *K_x, float *alpha_prime_x, float *a_x, float *b_x, float *K_x_half, float *alpha_prime_x_half,
                float *a_x_half, float *b_x_half, float *K_y, float *alpha_prime_y, float *a_y, float *b_y,
                float *K_y_half, float *alpha_prime_y_half, float *a_y_half, float *b_y_half, float *K_z,
                float *alpha_prime_z, float *a_z, float *b_z, float *K_z_half, float *alpha_prime_z_half,
                float *a_z_half, float *b_z_half)
{
    const ifos_geom &g = geom();
    const int len = 2 * FW + 1;
    std::vector<float> prof(18 * (size_t)len), alpha(6 * (size_t)len);
    ck(ifos_cpml_profiles(&g, VPPML, FPML, prof.data()), "CPML_coeff");
    ck(ifos_cpml_alpha(&g, FPML, alpha.data()), "CPML_coeff");
    float *kv[18] = {K_x, a_x, b_x, K_x_half, a_x_half, b_x_half, K_y, a_y, b_y,
                     K_y_half, a_y_half, b_y_half, K_z, a_z, b_z, K_z_half, a_z_half, b_z_half};
    float *av[6] = {alpha_prime_x, alpha_prime_x_half, alpha_prime_y, alpha_prime_y_half, alpha_prime_z,
                    alpha_prime_z_half};
    for (int n = 0; n < 18; ++n) memcpy(&kv[n][1], &prof[(size_t)n * len + 1], sizeof(float) * 2 * FW);
    for (int n = 0; n < 6; ++n) memcpy(&av[n][1], &alpha[(size_t)n * len + 1], sizeof(float) * 2 * FW);
    S.prof_dirty = true;
}

// CPML_ini_elastic.c:26-71 (one rank)
void CPML_ini_elastic(int *xb, int *yb, int *zb)
{
    xb[0] = FW + 1;
    xb[1] = NX - FW;
    yb[0] = FREE_SURF ? 1 : FW + 1;
    yb[1] = NY - FW;
    zb[0] = FW + 1;
    zb[1] = NZ - FW;
}

void zero_wavefield(int nx, int ny, int nz, float ***vx, float ***vy, float ***vz, float ***sxx, float ***syy,
                    float ***szz, float ***sxy, float ***syz, float ***sxz, float ***rxx, float ***ryy, float ***rzz,
                    float ***rxy, float ***ryz, float ***rxz, float ***psi_sxx_x, float ***psi_sxy_x,
                    float ***psi_sxz_x, float ***psi_sxy_y, float ***psi_syy_y, float ***psi_syz_y,
                    float ***psi_sxz_z, float ***psi_syz_z, float ***psi_szz_z, float ***psi_vxx, float ***psi_vyx,
                    float ***psi_vzx, float ***psi_vxy, float ***psi_vyy, float ***psi_vzy, float ***psi_vxz,
                    float ***psi_vyz, float ***psi_vzz)
{
    (void)nx, (void)ny, (void)nz;
    geom();
    if (L) {  // zero_wavefield.c:80-95
        float ***r[6] = {rxx, ryy, rzz, rxy, ryz, rxz};
        for (int n = 0; n < 6; ++n) ck(ifos_fill(m3(r[n], AVG, false), S.fe, 0.0f, nullptr), "zero_wavefield");
    }
    float ***w[9] = {vx, vy, vz, sxx, syy, szz, sxy, syz, sxz};
    const Kind wk[9] = {WAVE_TOP, WAVE_TOP, WAVE_TOP, WAVE, WAVE, WAVE, WAVE_TOP, WAVE_TOP, WAVE};
    for (int n = 0; n < 9; ++n) ck(ifos_fill(m3(w[n], wk[n], false), S.fe, 0.0f, nullptr), "zero_wavefield");
    vy[NY / 2][NX / 2][NZ / 2] = 0.0f;  // the host probe cell
    if (ABS_TYPE == 1) {
        float ***p[18] = {psi_sxx_x, psi_sxy_x, psi_sxz_x, psi_sxy_y, psi_syy_y, psi_syz_y, psi_sxz_z, psi_syz_z,
                          psi_szz_z, psi_vxx, psi_vyx, psi_vzx, psi_vxy, psi_vyy, psi_vzy, psi_vxz, psi_vyz, psi_vzz};
        for (int n = 0; n < 18; ++n) {
            const int m = n % 9;
            const Kind k = m < 3 ? PSI_X : m < 6 ? PSI_Y : PSI_Z;
            ck(ifos_fill(m3(p[n], k, false), psi_elems(k), 0.0f, nullptr), "zero_wavefield");
        }
    }
}

void zero_invers(int nx, int ny, int nz, float ****Ffvx, float ****Ffvy, float ****Ffvz, float ****Ffivx,
                 float ****Ffivy, float ****Ffivz, float ****Fbvx, float ****Fbvy, float ****Fbvz, float ****Fbivx,
                 float ****Fbivy, float ****Fbivz, int nfmax, int ntr_hess)
{
    (void)nx, (void)ny, (void)nz;
    geom();
    if (ntr_hess < 0) fatal("zero_invers: ntr_hess %d", ntr_hess);
    float ****F[12] = {Ffvx, Ffvy, Ffvz, Ffivx, Ffivy, Ffivz, Fbvx, Fbvy, Fbvz, Fbivx, Fbivy, Fbivz};
    // zero_invers.c:48-79: the adjoint side holds nfmax * (ntr_hess + 1) entries
    for (int n = 0; n < 12; ++n) {
        const int m = n < 6 ? nfmax : nfmax * (ntr_hess + 1);
        ck(ifos_fill(m4(F[n], m, false), S.fe * (size_t)m, 0.0f, nullptr), "zero_invers");
    }
}

void zero_grad(int nx, int ny, int nz, float ***grad1, float ***grad2, float ***grad3)
{
    (void)nx, (void)ny, (void)nz;
    geom();
    float ***G[3] = {grad1, grad2, grad3};
    for (int n = 0; n < 3; ++n) {
        ck(ifos_fill(m3(G[n], AVG, false), S.fe, 0.0f, nullptr), "zero_grad");
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i) memset(&G[n][j][i][1], 0, sizeof(float) * NZ);
    }
}

const char *ifos_dropin_last_error(void) { return S.err.c_str(); }

void ifos_dropin_release(void)
{
    flush_seis();
    if (S.lb) ifos_lbfgs_destroy(S.lb);
    S.lb = nullptr;
    S.lb_key = nullptr;
    (void)hipDeviceSynchronize();
    if (S.probe_host) {
        for (int q = 0; q < 2; ++q) (void)hipEventDestroy(S.probe_ev[q]);
        (void)hipHostFree(S.probe_host);
        S.probe_host = nullptr;
        S.probe_have[0] = S.probe_have[1] = false;
    }
    for (auto &kv : S.mir) (void)hipFree(kv.second.dev);
    S.mir.clear();
    void *bufs[] = {S.d_prof, S.d_src, S.d_rec, S.d_sig, S.d_adj, S.d_seisrec, S.d_step};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    S.d_step = nullptr;
    S.cap_step = 0;
    S.d_prof = nullptr;
    S.d_src = S.d_rec = S.d_chain = S.d_seisrec = nullptr;
    S.d_sig = S.d_adj = S.d_sec = nullptr;
    S.cap_src = S.cap_rec = S.cap_sig = S.cap_adj = S.cap_seis = 0;
    S.prof_key = nullptr;
    S.prof_dirty = true;
    S.src_n = S.src_back = -1;
    S.seis_key = nullptr;
    S.seis_ntr = -1;
    S.seis_ns = S.seis_pending = 0;
    for (int c = 0; c < 3; ++c) S.seis_out[c] = nullptr;
    S.have_geom = false;
}

}  // extern "C"
