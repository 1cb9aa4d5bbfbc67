// ifos_dev.h -- device-side geometry and helpers shared by the gfx950 kernels.
//
// All kernels keep the reference's arithmetic order and are compiled with
// -ffp-contract=off and IEEE division, so an FP32 update is bit-identical to
// the reference C code on the same inputs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ifos3d_hip.h"

// largest number of DFT frequencies per call (kernel-argument arrays)
#define IFOS_MAX_NF 32

namespace ifos {

// Geometry in the form the kernels use, passed by value.
struct DevGeom {
    int nx, ny, nz, fw;
    int free_surf;
    int64_t sj, si;      // element strides of j and i (k stride is 1)
    int64_t base;        // offset of (0,0,0)
    float dx, dy, dz, dt;
    float dtdx, dtdy, dtdz;  // DT/DX ... in float, update_v_ssg.c:95-97
    float b1, b2;            // 4th-order coefficients, update_v_ssg.c:98-100
    uint32_t fbytes;         // bytes of one device-layout field
    float rdx, rdy, rdz;     // 1/DX ... when p2
    int p2;                  // every spacing a power of two with a normal reciprocal
};

// The nine velocity derivatives xx xy xz yx yy yz zx zy zz (their 4th-order
// numerators) divided by the spacing of their second axis, as
// update_s_ssg_elastic.c:99-115 divides them.  When every spacing is 2^e with
// 2^-e a normal float, x * 2^-e is the same real number as x / 2^e and rounds
// to the same float (subnormal results included), so the multiply replaces
// the IEEE division bit for bit; otherwise the division is kept.
__device__ __forceinline__ void div_spacing(const DevGeom &g, float &xx, float &xy, float &xz, float &yx, float &yy,
                                            float &yz, float &zx, float &zy, float &zz)
{
    if (g.p2) {
        xx *= g.rdx; xy *= g.rdy; xz *= g.rdz;
        yx *= g.rdx; yy *= g.rdy; yz *= g.rdz;
        zx *= g.rdx; zy *= g.rdy; zz *= g.rdz;
    } else {
        xx = xx / g.dx; xy = xy / g.dy; xz = xz / g.dz;
        yx = yx / g.dx; yy = yy / g.dy; yz = yz / g.dz;
        zx = zx / g.dx; zy = zy / g.dy; zz = zz / g.dz;
    }
}

__host__ __device__ inline int64_t at(const DevGeom &g, int j, int i, int k)
{
    return g.base + (int64_t)j * g.sj + (int64_t)i * g.si + k;
}

// Slab index of a cell along one axis: [1..fw] -> h, [n-fw+1..n] -> h in
// [fw+1..2fw], else 0 (update_v_ssg_CPML.c:70-375 index maps).
__host__ __device__ inline int slab_h(int idx, int n, int fw, int lo_active)
{
    if (lo_active && idx <= fw) return idx;
    if (idx >= n - fw + 1) return idx - (n - 2 * fw);
    return 0;
}

__device__ inline int64_t psi_x_at(const DevGeom &g, int j, int h, int k)
{
    return ((int64_t)(j - 1) * (2 * g.fw) + (h - 1)) * g.nz + (k - 1);
}
__device__ inline int64_t psi_y_at(const DevGeom &g, int h, int i, int k)
{
    return ((int64_t)(h - 1) * g.nx + (i - 1)) * g.nz + (k - 1);
}
__device__ inline int64_t psi_z_at(const DevGeom &g, int j, int i, int h)
{
    return ((int64_t)(j - 1) * g.nx + (i - 1)) * (2 * g.fw) + (h - 1);
}

// CPML profile accessor: prof[(6*axis + n)*(2fw+1) + h], n = K a b Kh ah bh
struct Prof {
    const float *p;
    int len;  // 2fw+1
    __device__ float K(int ax, int h) const { return p[(6 * ax + 0) * len + h]; }
    __device__ float A(int ax, int h) const { return p[(6 * ax + 1) * len + h]; }
    __device__ float B(int ax, int h) const { return p[(6 * ax + 2) * len + h]; }
    __device__ float KH(int ax, int h) const { return p[(6 * ax + 3) * len + h]; }
    __device__ float AH(int ax, int h) const { return p[(6 * ax + 4) * len + h]; }
    __device__ float BH(int ax, int h) const { return p[(6 * ax + 5) * len + h]; }
};

// psi <- b*psi + a*d ;  d <- d/K + psi   (update_v_ssg_CPML.c:85-90)
__device__ inline void cpml_apply(float *psi, int64_t at_, float &d, float K, float a, float b)
{
    float p = b * psi[at_] + a * d;
    psi[at_] = p;
    d = d / K + p;
}

struct Psi9 {
    float *p[9];
};
// shots of one launch: shot s uses wavefield arrays at +s*fstride and CPML
// memory arrays at +s*pstride[axis] (elements); the model is shared
struct Shots {
    int n;
    int64_t fstride;
    int64_t pstride[3];
};
// velocity snapshot of the batched path: shot s, component n of the slot at
// ring + s*rstride + soff/4 + n*(field elements) + cell (ring points at the
// slot; rbytes covers the slot's three components)
struct Snap {
    float *ring;
    int64_t rstride;
    uint32_t rbytes;
    uint32_t soff;
};

// viscoelastic arrays of the stress march: memory variables r (per shot,
// wavefield layout), relaxation-time ratios (shared model), eta, L
// with L mechanisms each component's memory variables are L fields back to
// back (mechanism l at r[n] + l * lstride elements; a batch's shot s at
// r[n] + s * L * fstride); eta[l] = DT / relaxation time of mechanism l
struct ViscoF {
    float *r[6];
    const float *taus, *taup, *tausipjp, *tausjpkp, *tausipkp;
    float eta[IFOS_MAX_L];
    int L;
    int64_t lstride;
};

// Index box (inclusive), and a list of them for the slab launches.
struct Box {
    int i0, i1, j0, j1, k0, k1;
};
struct BoxList {
    Box b[6];
    int64_t start[7];  // cumulative cell counts
    int n;
};

// Tile shape of a march launch: TI i-rows per block (1, 2, 4, 8), jc
// j-planes per tile (<= 0: automatic), `group` consecutive tiles per shot
// before the next shot (1: shots fastest, <= 0: one shot after the other),
// split: interior box and CPML slabs in separate launches; pipe: the next
// plane's operands are loaded one iteration ahead (two planes in flight).
struct MarchCfg {
    int ti, jc, group;
    int split;
    int pipe;
    int kb;  // wavefronts of a workgroup side by side along k (divides ti; <= 1: one)
    int sb;  // shots of a workgroup, one wavefront row each (kb * sb divides ti)
    int lds; // in-plane neighbours staged in LDS (64-lane rows, kb = sb = 1, ti >= 2)
};
// TI / kb / sb checks of a tile shape {TI, jc, group, split, pipe, kb, sb, lds}
int check_tile_cfg(const int *c, const char *who);


// Raw buffer access: one 32-bit byte offset per thread, the neighbour
// displacements folded into the scalar offset and the instruction's
// immediate offset, so a load costs no vector address arithmetic.
// Word 3 of the descriptor is the gfx9-family value (raw, dword format).
typedef __amdgpu_buffer_rsrc_t rsrc_t;
__device__ inline rsrc_t make_rsrc(const void *p, uint32_t nbytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)nbytes, 0x00020000);
}
// the _b32 builtins move raw 32-bit words: bit casts, never value conversions
__device__ inline float bload(rsrc_t r, uint32_t voff, uint32_t soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
__device__ inline void bstore(float v, rsrc_t r, uint32_t voff, uint32_t soff)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)voff, (int)soff, 0);
}
// the same with a cache policy CP: IFOS_CPOL_NT is the non-temporal
// (streaming) hint, for words a kernel touches once, so that they do not
// evict the neighbour lines other workgroups still read from L2
constexpr int IFOS_CPOL_NT = 2;
template <int CP>
__device__ inline float bload_cp(rsrc_t r, uint32_t voff, uint32_t soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, CP));
}
template <int CP>
__device__ inline void bstore_cp(float v, rsrc_t r, uint32_t voff, uint32_t soff)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)voff, (int)soff, CP);
}

DevGeom make_dev_geom(const ifos_geom *g);

// multi-shot fused (interior + CPML) updates over [1..N]^3 for the batched
// path (stencil.hip): f and psi point at shot 0; snap may be null
int launch_v_shots(const ifos_geom *g, const ifos_fields *f, const ifos_model *m, const float *d_prof,
                   const ifos_psi *psi, const Shots &S, const Snap *snap, const MarchCfg &c, hipStream_t s);
int launch_s_shots(const ifos_geom *g, const ifos_fields *f, const ifos_model *m, const float *d_prof,
                   const ifos_psi *psi, const Shots &S, const MarchCfg &c, hipStream_t s,
                   const ViscoF *V = nullptr);
// gradient_F of ns shots in shot order into d_grad (aux.hip); shot s's DFT
// fields at d_F*[n] + s*fshot
int launch_gradient_shots(const ifos_geom *g, float *const d_Ff[6], float *const d_Fb[6], int64_t fshot, int ns,
                          const ifos_model *m, const float *finv, int nf, float *const d_grad[3], hipStream_t st);
// arguments of the fused stress(n) + velocity(n+1) step (stencil.hip)
struct FusedArgs {
    const float *vx, *vy, *vz, *sxx, *syy, *szz, *sxy, *syz, *sxz;  // v^n, sigma^(n-1)
    float *nvx, *nvy, *nvz, *nsxx, *nsyy, *nszz, *nsxy, *nsyz, *nsxz;  // v^(n+1), sigma^n
    const float *pi, *u, *uipjp, *ujpkp, *uipkp, *rip, *rjp, *rkp;
    Psi9 psv_old, psv_new;  // velocity-derivative memory: read, written by the owner
    Psi9 pss;               // stress-derivative memory, updated in place by the owner
    const int *src;         // [S][4] i j k type (null: no explosive source)
    const float *sig;       // [S][ldsig]
    int ldsig, nt, ntall;   // psource of step nt (psource.c:26-54)
};
struct FTiles {
    int ntk, nti, ntj, jc, ns, total, per;
};

int launch_vs_fused(const ifos_geom *g, const FusedArgs &A, const float *d_prof, const Shots &S, const Snap *snap,
                    int ti, int jc, hipStream_t s);

// gradient_F for one frequency of one shot, march kernel (stencil.hip); with
// hess, hess_F for one frequency of one receiver-side set
int launch_grad_march(const ifos_geom *g, const float *const f[12], float fd, const ifos_model *m,
                      float *const d_grad[3], hipStream_t s, bool hess = false);
// time-domain gradient (stencil.hip, aux.hip): one time sample of the
// lambda / mu / rho sums; u += v dt; the sums to the vp, vs, rho gradients
int launch_td_sample(const ifos_geom *g, const float *const uf[3], const float *const vf[3],
                     const float *const ub[3], const float *const vb[3], double w, double *const acc[3],
                     hipStream_t s);
int launch_td_integrate(const ifos_geom *g, float *const u[3], const float *const v[3], float dt, hipStream_t s);

// one reverse step n of the consistent adjoint (adjoint.hip)
struct CaStep {
    float *L[9], *scr[9];        // adjoint fields, scratch (device layout, zero halos)
    const float *v[3];           // forward v(n)
    const float *s[6];           // forward sigma(n-1): sxx syy szz sxy syz sxz
    float *pv[9], *ps[9];        // forward CPML memory after step n (read only)
    float *mu_v[9], *mu_s[9];    // their adjoints
    const float *model[9];       // rho pi u uipjp ujpkp uipkp rjp rkp rip
    double *G[8];                // dJ / d rip rjp rkp pi u uipjp ujpkp uipkp
    const int *rec;
    int ntr;
    const float *res;            // [3][ntr][nt+1]
    int n;
};
int launch_ca_step(const ifos_geom *g, const float *d_prof, const CaStep &S, hipStream_t st);
int launch_ca_residual(const float *d_syn, const float *d_obs, float *d_res, int ntr, int nt, double *d_J,
                       hipStream_t st);
int launch_ca_finish(const ifos_geom *g, double *const G[8], const ifos_model *m, double *Gr, double *Gp,
                     double *Gu, float *const d_grad[3], hipStream_t st);
int launch_td_finish(const ifos_geom *g, double *const acc[3], const ifos_model *m, float *const d_grad[3],
                     hipStream_t s);
void set_error(const char *fmt, ...);
int hip_check(hipError_t e, const char *what);

// Optional per-kernel timing (ifos_prof_enable): a pair of HIP events is
// recorded around each launch on the launch's own stream; prof_begin returns
// the pair's id (-1 when timing is off or the stream is being captured).
long prof_begin(int cat, hipStream_t s);
void prof_end(long id, hipStream_t s);
struct ProfScope {
    long id;
    hipStream_t s;
    ProfScope(int c, hipStream_t st) : id(prof_begin(c, st)), s(st) {}
    ~ProfScope() { prof_end(id, s); }
};

}  // namespace ifos
