// stencil.hip -- 4th-order staggered-grid velocity and stress updates for
// gfx950, elastic, with CPML memory variables updated only in the slabs.
//
// Two launch shapes:
//  * march kernels (k_v_march / k_s_march): a 64-wide wavefront spans k (the
//    contiguous axis, coalesced 256-B rows), a block spans TI rows of i, and
//    each thread marches a chunk of j planes keeping the j-direction operands
//    in a 4-deep register queue, so each j-neighbour is loaded from memory
//    once per column.  In-plane (i, k) neighbours come through L1/L2.
//    Blocks are renumbered so that the tiles each XCD receives form one
//    contiguous stretch of the grid (workgroups are dealt round-robin over
//    the 8 XCDs), which keeps the i-halo rows of a tile in the same L2.
//  * slab kernel (k_v_slabs / k_s_slabs): one thread per cell over the six
//    CPML slabs, the shape of update_v_ssg_CPML.c / update_s_ssg_CPML_elastic.c,
//    used by the reference-shaped (drop-in) update_*_CPML entry points.
//
// The march kernels also run several shots of one model in one launch (the
// batched path): logical tile L -> (tile, shot) with the shot fastest, so the
// workgroups of one XCD sweep the same tile for every shot back to back and
// the shared model arrays are read from HBM once per tile.  At the steps the
// DFT samples, the velocity kernel also writes the new velocities into a
// snapshot ring (the chunked DFT reads them later).
//
// Arithmetic follows the reference expression by expression
// (update_v_ssg.c:106-125, update_v_ssg_CPML.c:73-100,
//  update_s_ssg_elastic.c:106-138, update_s_ssg_CPML_elastic.c:63-140).
#include <algorithm>

#include "ifos_dev.h"

namespace ifos {

struct VFields {
    float *vx, *vy, *vz;
    const float *sxx, *syy, *szz, *sxy, *syz, *sxz;
    const float *rip, *rjp, *rkp;
};
struct SFields {
    const float *vx, *vy, *vz;
    float *sxx, *syy, *szz, *sxy, *syz, *sxz;
    const float *pi, *u, *uipjp, *ujpkp, *uipkp;
};
// 4th-order differences: forward f(c+1)-f(c), backward f(c)-f(c-1)
#define D4F(b1, b2, m1, c0, p1, p2) ((b1) * ((p1) - (c0)) + (b2) * ((p2) - (m1)))
#define D4B(b1, b2, m2, m1, c0, p1) ((b1) * ((c0) - (m1)) + (b2) * ((p1) - (m2)))

// ---------------------------------------------------------------------------
// velocity: CPML corrections of the 9 stress derivatives of one cell
// (update_v_ssg_CPML.c:84-130: x with K/a/b_x(_half), y, z)
__device__ inline void v_cpml(const DevGeom &g, const Prof &P, const Psi9 &psi, int j, int i, int k,
                              int hx, int hy, int hz, float &sxx_x, float &sxy_x, float &sxz_x,
                              float &sxy_y, float &syy_y, float &syz_y, float &sxz_z, float &syz_z,
                              float &szz_z)
{
    if (hx) {
        int64_t a = psi_x_at(g, j, hx, k);
        cpml_apply(psi.p[0], a, sxx_x, P.KH(0, hx), P.AH(0, hx), P.BH(0, hx));
        cpml_apply(psi.p[1], a, sxy_x, P.K(0, hx), P.A(0, hx), P.B(0, hx));
        cpml_apply(psi.p[2], a, sxz_x, P.K(0, hx), P.A(0, hx), P.B(0, hx));
    }
    if (hy) {
        int64_t a = psi_y_at(g, hy, i, k);
        cpml_apply(psi.p[3], a, sxy_y, P.K(1, hy), P.A(1, hy), P.B(1, hy));
        cpml_apply(psi.p[4], a, syy_y, P.KH(1, hy), P.AH(1, hy), P.BH(1, hy));
        cpml_apply(psi.p[5], a, syz_y, P.K(1, hy), P.A(1, hy), P.B(1, hy));
    }
    if (hz) {
        int64_t a = psi_z_at(g, j, i, hz);
        cpml_apply(psi.p[6], a, sxz_z, P.K(2, hz), P.A(2, hz), P.B(2, hz));
        cpml_apply(psi.p[7], a, syz_z, P.K(2, hz), P.A(2, hz), P.B(2, hz));
        cpml_apply(psi.p[8], a, szz_z, P.KH(2, hz), P.AH(2, hz), P.BH(2, hz));
    }
}

// stress: CPML corrections of the 9 velocity derivatives
// (update_s_ssg_CPML_elastic.c:72-115; the low-z layer divides by the y-axis
// K profile, :104-108)
__device__ inline void s_cpml(const DevGeom &g, const Prof &P, const Psi9 &psi, int j, int i, int k,
                              int hx, int hy, int hz, float &vxx, float &vyx, float &vzx, float &vxy,
                              float &vyy, float &vzy, float &vxz, float &vyz, float &vzz)
{
    if (hx) {
        int64_t a = psi_x_at(g, j, hx, k);
        cpml_apply(psi.p[0], a, vxx, P.K(0, hx), P.A(0, hx), P.B(0, hx));
        cpml_apply(psi.p[1], a, vyx, P.KH(0, hx), P.AH(0, hx), P.BH(0, hx));
        cpml_apply(psi.p[2], a, vzx, P.KH(0, hx), P.AH(0, hx), P.BH(0, hx));
    }
    if (hy) {
        int64_t a = psi_y_at(g, hy, i, k);
        cpml_apply(psi.p[3], a, vxy, P.KH(1, hy), P.AH(1, hy), P.BH(1, hy));
        cpml_apply(psi.p[4], a, vyy, P.K(1, hy), P.A(1, hy), P.B(1, hy));
        cpml_apply(psi.p[5], a, vzy, P.KH(1, hy), P.AH(1, hy), P.BH(1, hy));
    }
    if (hz) {
        int64_t a = psi_z_at(g, j, i, hz);
        const int kax = (k <= g.fw) ? 1 : 2;
        cpml_apply(psi.p[6], a, vxz, P.KH(kax, hz), P.AH(2, hz), P.BH(2, hz));
        cpml_apply(psi.p[7], a, vyz, P.KH(kax, hz), P.AH(2, hz), P.BH(2, hz));
        cpml_apply(psi.p[8], a, vzz, P.K(kax, hz), P.A(2, hz), P.B(2, hz));
    }
}

// stress update of one cell from its 9 velocity derivatives
// (update_s_ssg_elastic.c:117-138)
__device__ inline void s_apply(const DevGeom &g, const SFields &F, int64_t c, float vxx, float vxy,
                               float vxz, float vyx, float vyy, float vyz, float vzx, float vzy, float vzz)
{
    const float DT = g.dt;
    float fipjp = F.uipjp[c] * DT;
    float fjpkp = F.ujpkp[c] * DT;
    float fipkp = F.uipkp[c] * DT;
    float gg = F.pi[c];
    float f = 2.0f * F.u[c];
    float vxyyx = vxy + vyx;
    float vyzzy = vyz + vzy;
    float vxzzx = vxz + vzx;
    float vxxyyzz = vxx + vyy + vzz;
    float vyyzz = vyy + vzz;
    float vxxzz = vxx + vzz;
    float vxxyy = vxx + vyy;
    F.sxy[c] += (fipjp * vxyyx);
    F.syz[c] += (fjpkp * vyzzy);
    F.sxz[c] += (fipkp * vxzzx);
    F.sxx[c] += DT * ((gg * vxxyyzz) - (f * vyyzz));
    F.syy[c] += DT * ((gg * vxxyyzz) - (f * vxxzz));
    F.szz[c] += DT * ((gg * vxxyyzz) - (f * vxxyy));
}

// update_s_ssg.c:174-238: the viscoelastic stress step with one relaxation
// mechanism.  s = sxx syy szz sxy syz sxz and r = rxx ryy rzz rxy ryz rxz
// (old values in, new values out).  The double subexpressions of the
// reference (1.0 + L*tau, 2.0*u, eta*0.5) are evaluated in double and
// rounded to float on assignment, as the C code does.
__device__ inline void stress_visco(float DT, float eta, int L, float pi, float u, float uipjp, float ujpkp,
                                    float uipkp, float ts, float tp, float tsipjp, float tsjpkp, float tsipkp,
                                    float vxyyx, float vyzzy, float vxzzx, float vxxyyzz, float vyyzz,
                                    float vxxzz, float vxxyy, float s[6], float r[6])
{
    const float Lf = (float)L;
    const float dthalbe = (float)((double)DT / 2.0);
    const float fipjp = (float)((double)(uipjp * DT) * (1.0 + (double)(Lf * tsipjp)));
    const float fjpkp = (float)((double)(ujpkp * DT) * (1.0 + (double)(Lf * tsjpkp)));
    const float fipkp = (float)((double)(uipkp * DT) * (1.0 + (double)(Lf * tsipkp)));
    const float gg = (float)((double)pi * (1.0 + (double)(Lf * tp)));
    const float f = (float)((2.0 * (double)u) * (1.0 + (double)(Lf * ts)));
    s[3] = s[3] + ((fipjp * vxyyx) + (dthalbe * r[3]));
    s[4] = s[4] + ((fjpkp * vyzzy) + (dthalbe * r[4]));
    s[5] = s[5] + ((fipkp * vxzzx) + (dthalbe * r[5]));
    s[0] = s[0] + (DT * ((gg * vxxyyzz) - (f * vyyzz)) + (dthalbe * r[0]));
    s[1] = s[1] + (DT * ((gg * vxxyyzz) - (f * vxxzz)) + (dthalbe * r[1]));
    s[2] = s[2] + (DT * ((gg * vxxyyzz) - (f * vxxyy)) + (dthalbe * r[2]));
    const float b = (float)(1.0 / (1.0 + ((double)eta * 0.5)));
    const float cc = (float)(1.0 - ((double)eta * 0.5));
    const float dipjp = uipjp * eta * tsipjp, djpkp = ujpkp * eta * tsjpkp, dipkp = uipkp * eta * tsipkp;
    const float dd = (float)(((2.0 * (double)u) * (double)eta) * (double)ts);
    const float e = pi * eta * tp;
    r[3] = b * ((r[3] * cc) - (dipjp * vxyyx));
    r[4] = b * ((r[4] * cc) - (djpkp * vyzzy));
    r[5] = b * ((r[5] * cc) - (dipkp * vxzzx));
    r[0] = b * (((r[0] * cc) - (e * vxxyyzz)) + (dd * vyyzz));
    r[1] = b * (((r[1] * cc) - (e * vxxyyzz)) + (dd * vxxzz));
    r[2] = b * (((r[2] * cc) - (e * vxxyyzz)) + (dd * vxxyy));
#pragma unroll
    for (int n = 0; n < 6; ++n) s[n] = s[n] + (dthalbe * r[n]);
}

// the memory-variable recursion of one mechanism (update_s_ssg.c:211-230)
__device__ inline void visco_mech(float eta, float pi, float u, float uipjp, float ujpkp, float uipkp, float ts,
                                  float tp, float tsipjp, float tsjpkp, float tsipkp, float vxyyx, float vyzzy,
                                  float vxzzx, float vxxyyzz, float vyyzz, float vxxzz, float vxxyy, float r[6])
{
    const float b = (float)(1.0 / (1.0 + ((double)eta * 0.5)));
    const float cc = (float)(1.0 - ((double)eta * 0.5));
    const float dipjp = uipjp * eta * tsipjp, djpkp = ujpkp * eta * tsjpkp, dipkp = uipkp * eta * tsipkp;
    const float dd = (float)(((2.0 * (double)u) * (double)eta) * (double)ts);
    const float e = pi * eta * tp;
    r[3] = b * ((r[3] * cc) - (dipjp * vxyyx));
    r[4] = b * ((r[4] * cc) - (djpkp * vyzzy));
    r[5] = b * ((r[5] * cc) - (dipkp * vxzzx));
    r[0] = b * (((r[0] * cc) - (e * vxxyyzz)) + (dd * vyyzz));
    r[1] = b * (((r[1] * cc) - (e * vxxyyzz)) + (dd * vxxzz));
    r[2] = b * (((r[2] * cc) - (e * vxxyyzz)) + (dd * vxxyy));
}

// the viscoelastic stress step with V.L >= 1 mechanisms: where the one
// mechanism of update_s_ssg.c:174-238 enters, the L memory variables of the
// component are summed in mechanism order (the tau-method's SLS sum), before
// and after each mechanism is updated with its own eta.  r0 holds mechanism
// 0 (in registers, updated in place); ld(l, n) / st(l, n, v) read and write
// component n of mechanism l >= 1 at the cell.
template <class LDF, class STF>
__device__ inline void stress_visco_n(float DT, const ViscoF &V, float pi, float u, float uipjp, float ujpkp,
                                      float uipkp, float ts, float tp, float tsipjp, float tsjpkp, float tsipkp,
                                      float vxyyx, float vyzzy, float vxzzx, float vxxyyzz, float vyyzz,
                                      float vxxzz, float vxxyy, float s[6], float r0[6], LDF ld, STF st)
{
    const int L = V.L;
    const float Lf = (float)L;
    const float dthalbe = (float)((double)DT / 2.0);
    const float fipjp = (float)((double)(uipjp * DT) * (1.0 + (double)(Lf * tsipjp)));
    const float fjpkp = (float)((double)(ujpkp * DT) * (1.0 + (double)(Lf * tsjpkp)));
    const float fipkp = (float)((double)(uipkp * DT) * (1.0 + (double)(Lf * tsipkp)));
    const float gg = (float)((double)pi * (1.0 + (double)(Lf * tp)));
    const float f = (float)((2.0 * (double)u) * (1.0 + (double)(Lf * ts)));
    float sr[6];
#pragma unroll
    for (int n = 0; n < 6; ++n) sr[n] = r0[n];
    for (int l = 1; l < L; ++l)
#pragma unroll
        for (int n = 0; n < 6; ++n) sr[n] = sr[n] + ld(l, n);
    s[3] = s[3] + ((fipjp * vxyyx) + (dthalbe * sr[3]));
    s[4] = s[4] + ((fjpkp * vyzzy) + (dthalbe * sr[4]));
    s[5] = s[5] + ((fipkp * vxzzx) + (dthalbe * sr[5]));
    s[0] = s[0] + (DT * ((gg * vxxyyzz) - (f * vyyzz)) + (dthalbe * sr[0]));
    s[1] = s[1] + (DT * ((gg * vxxyyzz) - (f * vxxzz)) + (dthalbe * sr[1]));
    s[2] = s[2] + (DT * ((gg * vxxyyzz) - (f * vxxyy)) + (dthalbe * sr[2]));
    visco_mech(V.eta[0], pi, u, uipjp, ujpkp, uipkp, ts, tp, tsipjp, tsjpkp, tsipkp, vxyyx, vyzzy, vxzzx, vxxyyzz,
               vyyzz, vxxzz, vxxyy, r0);
#pragma unroll
    for (int n = 0; n < 6; ++n) sr[n] = r0[n];
    for (int l = 1; l < L; ++l) {
        float rl[6];
#pragma unroll
        for (int n = 0; n < 6; ++n) rl[n] = ld(l, n);
        visco_mech(V.eta[l], pi, u, uipjp, ujpkp, uipkp, ts, tp, tsipjp, tsjpkp, tsipkp, vxyyx, vyzzy, vxzzx,
                   vxxyyzz, vyyzz, vxxzz, vxxyy, rl);
#pragma unroll
        for (int n = 0; n < 6; ++n) {
            st(l, n, rl[n]);
            sr[n] = sr[n] + rl[n];
        }
    }
#pragma unroll
    for (int n = 0; n < 6; ++n) s[n] = s[n] + (dthalbe * sr[n]);
}

// ---------------------------------------------------------------------------
// Tiles.  A launch covers up to six boxes (the interior box, the CPML slabs,
// or all of [1..N]^3), each cut into tiles of kw k-columns x (TI * rpw)
// i-rows x jc j-planes: a wavefront holds rpw = 64 / kw rows of kw lanes, so
// a thin box (a 10-cell z slab) still fills most lanes.  Logical tile L,
// for ns shots in groups of `group` consecutive tiles per shot, is
// renumbered XCD-aware: physical block b runs on XCD b % 8, and
// L = (b % 8) * per + b / 8 gives each XCD a contiguous range.
struct MTiles {
    Box b[6];
    int kw[6], rpw[6], ntk[6], nti[6], kb[6];
    int start[7];  // cumulative tile counts per box
    int nb, jc, ns, group, total, per;
    int sb, nsg;  // shots per workgroup (one wavefront row each), shot groups
};

__device__ inline bool decode_tile(const MTiles &T, int TI, int &k, int &i, int &jb, int &je, int &shot,
                                   Box &R, bool *krow = nullptr)
{
    const int b = blockIdx.x;
    const int L = (b & 7) * T.per + (b >> 3);
    if (L >= T.total) return false;
    const int gs = T.group * T.nsg;
    const int rem = L % gs;
    const int sgrp = rem / T.group;
    int t = (L / gs) * T.group + rem % T.group;
    if (t >= T.start[T.nb]) return false;
    int q = 0;
    while (q + 1 < T.nb && t >= T.start[q + 1]) ++q;
    t -= T.start[q];
    R = T.b[q];
    const int kw = T.kw[q], rpw = T.rpw[q], kb = T.kb[q];
    const int dk = (int)threadIdx.x % kw, di = (int)threadIdx.x / kw;
    if (di >= rpw) return false;
    const int tk = t % T.ntk[q];
    t /= T.ntk[q];
    const int ti = t % T.nti[q], tj = t / T.nti[q];
    // wavefront y of the workgroup: k column y % kb, shot (y / kb) % sb of
    // the shot group, i row y / (kb sb)
    const int y = (int)threadIdx.y, sb = T.sb;
    const int wk = y % kb, ws = (y / kb) % sb, wi = y / (kb * sb);
    // one value per wavefront (blockDim.x is the wavefront): keep the shot,
    // and with it the field descriptors, in scalar registers
    shot = __builtin_amdgcn_readfirstlane(sgrp * sb + ws);
    if (shot >= T.ns) return false;
    k = R.k0 + (tk * kb + wk) * kw + dk;
    i = R.i0 + (ti * (TI / (kb * sb)) + wi) * rpw + di;
    if (k > R.k1 || i > R.i1) return false;
    jb = R.j0 + tj * T.jc;
    je = min(jb + T.jc - 1, R.j1);
    // the wavefront spans 64 consecutive k of one (j, i) row, all inside the box
    if (krow) *krow = __builtin_amdgcn_readfirstlane((int)(kw == 64 && k - dk + 63 <= R.k1)) != 0;
    return true;
}

// k-neighbours through the lanes of a wavefront that spans 64 consecutive k
// of one row (DPP wavefront shifts instead of loads): lane l holds the
// row's value c at k0 + l; h holds the values just outside the row in lanes
// 0, 1 (k0 - 2, k0 - 1) and 62, 63 (k0 + 64, k0 + 65), loaded with the
// per-lane offset khalo_off.  Lanes whose shift source lies outside the
// wavefront keep the `old` operand of the DPP move.
template <int CTRL>
__device__ inline float dpp_move(float old, float src)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src), CTRL, 0xF, 0xF, false));
}
constexpr int DPP_WAVE_SHL1 = 0x130, DPP_WAVE_SHR1 = 0x138;
struct KNb {
    float m2, m1, p1, p2;  // values at k-2, k-1, k+1, k+2
};
__device__ inline KNb kshift(float c, float h)
{
    KNb r;
    const float hs = dpp_move<DPP_WAVE_SHL1>(h, h);  // lane 0: h[1]
    const float hr = dpp_move<DPP_WAVE_SHR1>(h, h);  // lane 63: h[62]
    r.m1 = dpp_move<DPP_WAVE_SHR1>(hs, c);
    r.m2 = dpp_move<DPP_WAVE_SHR1>(h, r.m1);
    r.p1 = dpp_move<DPP_WAVE_SHL1>(hr, c);
    r.p2 = dpp_move<DPP_WAVE_SHL1>(h, r.p1);
    return r;
}
// byte offset (within the LD addressing, where 4 x is the word at k - 2 + x)
// of lane l's halo element, k = k0 + l: k0 - 2 and k0 - 1 for lanes 0, 1,
// k0 + 64 and k0 + 65 for lanes 62, 63.  kshift reads no other lane of h, so
// lanes 2-31 load lane 1's word and lanes 32-61 lane 62's: the instruction
// touches only the two lines just outside the row, not the row's own lines
// again.  Lanes 2-31 step back (1 - l words): vo + offset wraps as uint32 to
// vo - 4 (l - 1), still positive, as vo counts at least k + IFOS_KOFF - 2 >=
// l + IFOS_KOFF - 1 words (k >= k0 + l, k0 >= 1).
__device__ inline uint32_t khalo_off()
{
    const int l = (int)threadIdx.x;
    const int x = l < 2 ? 0 : l < 32 ? 1 - l : l < 63 ? 66 - l : 4;
    return (uint32_t)(x * 4);
}

__device__ inline void shift_psi(Psi9 &psi, const Shots &S, int shot)
{
#pragma unroll
    for (int n = 0; n < 9; ++n) psi.p[n] += (int64_t)shot * S.pstride[n / 3];
}

// Byte offsets for the march kernels: vo is the offset of (j-2, i-2, k-2)
// of the thread's column at the current plane j, so every operand
// (j+dj, i+di, k+dk), |d| <= 2 (dj <= 3), lies at vo + (dk+2)*4 (immediate)
// + (dj+2)*SJ + (di+2)*SI (scalar): all non-negative, all uniform but vo.
#define LD(r, dj, di, dk) bload(r, vo + ((dk) + 2) * 4, ((dj) + 2) * SJ + ((di) + 2) * SI)
#define LDH(r, dj) bload(r, vo + khv, ((dj) + 2) * SJ + 2 * SI)
// words a march touches once (its own cell's old value, its stores) carry
// the policy CPN, the streaming hint in the whole-domain launches (CPML
// template true: the batched path, whose fields are far larger than the
// 256 MB Infinity Cache) and the default policy in the interior launches of
// the plain path (measured: streaming there makes them 10-25 % slower, as
// the next launches re-read those words from the Infinity Cache)
#define LDN(r, dj, di, dk) bload_cp<CPN>(r, vo + ((dk) + 2) * 4, ((dj) + 2) * SJ + ((di) + 2) * SI)
#define STN(v, r) bstore_cp<CPN>(v, r, vo + 8, 2 * SJ + 2 * SI)
#define ST(v, r) bstore(v, r, vo + 8, 2 * SJ + 2 * SI)

// Operands of one plane j of the velocity march: the in-plane neighbours,
// the model, the old velocities, and the j-queue values first needed at j
// (sxy, syz at j+1, syy at j+2).  D = 0 loads plane j of the current offset,
// D = 1 the next plane (prefetched one iteration ahead when PIPE).
struct VPlane {
    float sxx_a, sxx_b, sxx_c, sxx_d, sxy_a, sxy_b, sxy_d, sxz_xa, sxz_xb, sxz_0, sxz_xd, sxz_za, sxz_zb, sxz_zd,
        syz_za, syz_zb, syz_zd, szz_a, szz_b, szz_c, szz_d, rip, rjp, rkp, vx0, vy0, vz0, sxy_n, syz_n, syy_n;
    float sxz_h, syz_h, szz_h;  // k-row halos (krow wavefronts)
};
#define VLOAD(P, D)                                                                                            \
    do {                                                                                                       \
        P.sxx_a = LD(rsxx, D, -1, 0); P.sxx_b = LD(rsxx, D, 0, 0); P.sxx_c = LD(rsxx, D, 1, 0);                  \
        P.sxx_d = LD(rsxx, D, 2, 0);                                                                           \
        P.sxy_a = LD(rsxy, D, -2, 0); P.sxy_b = LD(rsxy, D, -1, 0); P.sxy_d = LD(rsxy, D, 1, 0);                 \
        P.sxz_xa = LD(rsxz, D, -2, 0); P.sxz_xb = LD(rsxz, D, -1, 0); P.sxz_0 = LD(rsxz, D, 0, 0);               \
        P.sxz_xd = LD(rsxz, D, 1, 0);                                                                          \
        if (krow) {                                                                                            \
            P.sxz_h = LDH(rsxz, D); P.syz_h = LDH(rsyz, D); P.szz_h = LDH(rszz, D);                            \
            P.szz_b = LD(rszz, D, 0, 0);                                                                       \
        } else {                                                                                               \
            P.sxz_za = LD(rsxz, D, 0, -2); P.sxz_zb = LD(rsxz, D, 0, -1); P.sxz_zd = LD(rsxz, D, 0, 1);          \
            P.syz_za = LD(rsyz, D, 0, -2); P.syz_zb = LD(rsyz, D, 0, -1); P.syz_zd = LD(rsyz, D, 0, 1);          \
            P.szz_a = LD(rszz, D, 0, -1); P.szz_b = LD(rszz, D, 0, 0); P.szz_c = LD(rszz, D, 0, 1);              \
            P.szz_d = LD(rszz, D, 0, 2);                                                                       \
        }                                                                                                      \
        P.rip = LD(rrip, D, 0, 0); P.rjp = LD(rrjp, D, 0, 0); P.rkp = LD(rrkp, D, 0, 0);                         \
        P.vx0 = LDN(rvx, D, 0, 0); P.vy0 = LDN(rvy, D, 0, 0); P.vz0 = LDN(rvz, D, 0, 0);                         \
        P.sxy_n = LD(rsxy, (D) + 1, 0, 0); P.syz_n = LD(rsyz, (D) + 1, 0, 0); P.syy_n = LD(rsyy, (D) + 2, 0, 0); \
    } while (0)

template <int TI, bool CPML, bool SNAP, bool PIPE>
__global__ __launch_bounds__(64 * TI) void k_v_march(DevGeom g, MTiles T, VFields F, Prof P, Psi9 psi, Shots S,
                                                    Snap N)
{
    int k, i, jb, je, shot;
    Box R;
    bool krow;
    if (!decode_tile(T, TI, k, i, jb, je, shot, R, &krow)) return;
    const uint32_t khv = khalo_off();
    constexpr int CPN = CPML ? IFOS_CPOL_NT : 0;
    const uint32_t SJ = (uint32_t)g.sj * 4u, SI = (uint32_t)g.si * 4u;
    const float b1 = g.b1, b2 = g.b2;
    const int64_t so = (int64_t)shot * S.fstride;
    const rsrc_t rsxx = make_rsrc(F.sxx + so, g.fbytes), rsyy = make_rsrc(F.syy + so, g.fbytes),
                 rszz = make_rsrc(F.szz + so, g.fbytes), rsxy = make_rsrc(F.sxy + so, g.fbytes),
                 rsyz = make_rsrc(F.syz + so, g.fbytes), rsxz = make_rsrc(F.sxz + so, g.fbytes),
                 rvx = make_rsrc(F.vx + so, g.fbytes), rvy = make_rsrc(F.vy + so, g.fbytes),
                 rvz = make_rsrc(F.vz + so, g.fbytes), rrip = make_rsrc(F.rip, g.fbytes),
                 rrjp = make_rsrc(F.rjp, g.fbytes), rrkp = make_rsrc(F.rkp, g.fbytes);
    rsrc_t rring;
    if (SNAP) rring = make_rsrc(N.ring + (int64_t)shot * N.rstride, N.rbytes);
    if (CPML) shift_psi(psi, S, shot);
    uint32_t vo = (uint32_t)((int64_t)jb * g.sj + (int64_t)i * g.si + k + IFOS_KOFF - 2) * 4u;
    int hx = 0;
    if (CPML) hx = slab_h(i, g.nx, g.fw, 1);
    const int hz = CPML ? slab_h(k, g.nz, g.fw, 1) : 0;

    // j queues: sxy and syz at j-2..j, syy at j-1..j+1 (the next element of
    // each arrives with the plane's operands)
    float sxy_m2 = LD(rsxy, -2, 0, 0), sxy_m1 = LD(rsxy, -1, 0, 0), sxy_0 = LD(rsxy, 0, 0, 0);
    float syz_m2 = LD(rsyz, -2, 0, 0), syz_m1 = LD(rsyz, -1, 0, 0), syz_0 = LD(rsyz, 0, 0, 0);
    float syy_m1 = LD(rsyy, -1, 0, 0), syy_0 = LD(rsyy, 0, 0, 0), syy_p1 = LD(rsyy, 1, 0, 0);
    VPlane cur, nxt;
    VLOAD(cur, 0);

    for (int j = jb; j <= je; ++j, vo += SJ) {
        if (PIPE && j < je) VLOAD(nxt, 1);
        if (krow) {
            const KNb a = kshift(cur.sxz_0, cur.sxz_h), b = kshift(syz_0, cur.syz_h), c = kshift(cur.szz_b, cur.szz_h);
            cur.sxz_za = a.m2; cur.sxz_zb = a.m1; cur.sxz_zd = a.p1;
            cur.syz_za = b.m2; cur.syz_zb = b.m1; cur.syz_zd = b.p1;
            cur.szz_a = c.m1; cur.szz_c = c.p1; cur.szz_d = c.p2;
        }
        float sxx_x = g.dtdx * D4F(b1, b2, cur.sxx_a, cur.sxx_b, cur.sxx_c, cur.sxx_d);
        float sxy_y = g.dtdy * D4B(b1, b2, sxy_m2, sxy_m1, sxy_0, cur.sxy_n);
        float sxz_z = g.dtdz * D4B(b1, b2, cur.sxz_za, cur.sxz_zb, cur.sxz_0, cur.sxz_zd);
        float syy_y = g.dtdy * D4F(b1, b2, syy_m1, syy_0, syy_p1, cur.syy_n);
        float sxy_x = g.dtdx * D4B(b1, b2, cur.sxy_a, cur.sxy_b, sxy_0, cur.sxy_d);
        float syz_z = g.dtdz * D4B(b1, b2, cur.syz_za, cur.syz_zb, syz_0, cur.syz_zd);
        float szz_z = g.dtdz * D4F(b1, b2, cur.szz_a, cur.szz_b, cur.szz_c, cur.szz_d);
        float sxz_x = g.dtdx * D4B(b1, b2, cur.sxz_xa, cur.sxz_xb, cur.sxz_0, cur.sxz_xd);
        float syz_y = g.dtdy * D4B(b1, b2, syz_m2, syz_m1, syz_0, cur.syz_n);
        if (CPML) {
            const int hy = slab_h(j, g.ny, g.fw, !g.free_surf);
            v_cpml(g, P, psi, j, i, k, hx, hy, hz, sxx_x, sxy_x, sxz_x, sxy_y, syy_y, syz_y, sxz_z, syz_z,
                   szz_z);
        }
        const float nvx = cur.vx0 + (sxx_x + sxy_y + sxz_z) / cur.rip;
        const float nvy = cur.vy0 + (syy_y + sxy_x + syz_z) / cur.rjp;
        const float nvz = cur.vz0 + (szz_z + sxz_x + syz_y) / cur.rkp;
        sxy_m2 = sxy_m1; sxy_m1 = sxy_0; sxy_0 = cur.sxy_n;
        syz_m2 = syz_m1; syz_m1 = syz_0; syz_0 = cur.syz_n;
        syy_m1 = syy_0; syy_0 = syy_p1; syy_p1 = cur.syy_n;
        if (!PIPE && j < je) VLOAD(nxt, 1);  // next plane's operands, issued before this plane's stores
        STN(nvx, rvx);
        STN(nvy, rvy);
        STN(nvz, rvz);
        if (SNAP) {
            bstore_cp<CPN>(nvx, rring, vo + 8, 2 * SJ + 2 * SI + N.soff);
            bstore_cp<CPN>(nvy, rring, vo + 8, 2 * SJ + 2 * SI + N.soff + g.fbytes);
            bstore_cp<CPN>(nvz, rring, vo + 8, 2 * SJ + 2 * SI + N.soff + 2 * g.fbytes);
        }
        cur = nxt;
    }
}
#undef VLOAD

// Velocity march with the in-plane neighbours staged in LDS, built like
// k_s_lds (which see for the staging, the clamps and the pipelining): the
// five stresses a velocity update differentiates along i or k go through
// LDS -- sxx, sxy and sxz for their i-neighbours (rows just outside the
// workgroup: sxx at i0-1, i0+TI, i0+TI+1; sxy and sxz at i0-2, i0-1,
// i0+TI), sxz, syz and szz for their k-neighbours (the row's halo words).
// sxx, sxz and szz are read at the plane itself, sxy and syz come from the
// j-queues.  Same arithmetic, in the same order, as k_v_march.
template <int TI, bool CPML, bool SNAP>
__global__ __launch_bounds__(64 * TI) void k_v_lds(DevGeom g, MTiles T, VFields F, Prof P, Psi9 psi, Shots S,
                                                  Snap N)
{
    constexpr int RW = TI + 4;
    constexpr int NH = (9 + TI - 1) / TI;
    constexpr int CPN = CPML ? IFOS_CPOL_NT : 0;
    __shared__ float sh[2][5][RW][68];  // sxx, sxy, sxz, syz, szz
    const int b = blockIdx.x;
    const int L = (b & 7) * T.per + (b >> 3);
    if (L >= T.total) return;  // whole workgroup
    const int gs = T.group * T.nsg;
    const int rem = L % gs;
    const int shot = rem / T.group;
    int t = (L / gs) * T.group + rem % T.group;
    if (t >= T.start[T.nb] || shot >= T.ns) return;  // whole workgroup
    int q = 0;
    while (q + 1 < T.nb && t >= T.start[q + 1]) ++q;
    t -= T.start[q];
    const Box R = T.b[q];
    const int tk = t % T.ntk[q];
    t /= T.ntk[q];
    const int ti = t % T.nti[q], tj = t / T.nti[q];
    const int lane = (int)threadIdx.x;
    const int y = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int k0 = R.k0 + tk * 64, i0 = R.i0 + ti * TI;
    const int k = k0 + lane, i = i0 + y;
    const int jb = R.j0 + tj * T.jc, je = min(jb + T.jc - 1, R.j1);
    const bool act = k <= R.k1 && i <= R.i1;
    const int kc = min(k, g.nz + 2);
    const int yc = min(i, g.nx + 2) - i0;
    const uint32_t SJ = (uint32_t)g.sj * 4u, SI = (uint32_t)g.si * 4u;
    const float b1 = g.b1, b2 = g.b2;
    const int64_t so = (int64_t)shot * S.fstride;
    const rsrc_t rsxx = make_rsrc(F.sxx + so, g.fbytes), rsyy = make_rsrc(F.syy + so, g.fbytes),
                 rszz = make_rsrc(F.szz + so, g.fbytes), rsxy = make_rsrc(F.sxy + so, g.fbytes),
                 rsyz = make_rsrc(F.syz + so, g.fbytes), rsxz = make_rsrc(F.sxz + so, g.fbytes),
                 rvx = make_rsrc(F.vx + so, g.fbytes), rvy = make_rsrc(F.vy + so, g.fbytes),
                 rvz = make_rsrc(F.vz + so, g.fbytes), rrip = make_rsrc(F.rip, g.fbytes),
                 rrjp = make_rsrc(F.rjp, g.fbytes), rrkp = make_rsrc(F.rkp, g.fbytes);
    rsrc_t rring;
    if (SNAP) rring = make_rsrc(N.ring + (int64_t)shot * N.rstride, N.rbytes);
    if (CPML) shift_psi(psi, S, shot);
    uint32_t vb = (uint32_t)((int64_t)jb * g.sj + (int64_t)i0 * g.si + kc + IFOS_KOFF - 2) * 4u;
    const uint32_t own = (uint32_t)(yc + 2) * SI;
#define LB(r, dj) bload(r, vb + 8u, (uint32_t)((dj) + 2) * SJ + own)
#define LBN(r, dj) bload_cp<CPN>(r, vb + 8u, (uint32_t)((dj) + 2) * SJ + own)
    const int kh = lane < 2 ? k0 - 2 + lane : lane < 32 ? k0 - 1 : lane < 63 ? k0 + 64 : k0 + 65;
    const int hcol = lane < 62 ? (lane < 2 ? lane : lane + 2) : lane + 4;
    const uint32_t hd = (uint32_t)((min(kh, g.nz + 2) - kc) * 4);
    // outside rows: e = min(y + n TI, 8) is (stress e / 3: sxx, sxy, sxz; row i0 + orow)
    int oc[NH], orow[NH];
    uint32_t ooff[NH];
#pragma unroll
    for (int n = 0; n < NH; ++n) {
        const int e = min(y + n * TI, 8);
        const int c = e < 3 ? 0 : (e < 6 ? 1 : 2), m = e - 3 * c;
        const int r = c == 0 ? (m == 0 ? -1 : (m == 1 ? TI : TI + 1)) : (m == 0 ? -2 : (m == 1 ? -1 : TI));
        oc[n] = c;
        orow[n] = r;
        ooff[n] = (uint32_t)(min(i0 + r, g.nx + 2) - i0 + 2) * SI;
    }
    int hx = 0;
    if (CPML) hx = slab_h(i, g.nx, g.fw, 1);
    const int hz = CPML ? slab_h(k, g.nz, g.fw, 1) : 0;

    struct Ops {
        float sxx, sxz, szz, rip, rjp, rkp, vx0, vy0, vz0;
        float hk[3], ho[NH];  // halo words of sxz, syz, szz; outside rows
    };
    // the cell's own operands, loaded at the top of their own iteration (as
    // in k_s_lds: prefetching them a plane ahead costs registers, not time)
    auto load_cell = [&](Ops &O, int D) {
        O.rip = LB(rrip, D); O.rjp = LB(rrjp, D); O.rkp = LB(rrkp, D);
        O.vx0 = LBN(rvx, D); O.vy0 = LBN(rvy, D); O.vz0 = LBN(rvz, D);
    };
    // what plane j + D stages in LDS, loaded a plane ahead
    auto load_ops = [&](Ops &O, int D) {
        const uint32_t pj = (uint32_t)(D + 2) * SJ;
        O.sxx = LB(rsxx, D); O.sxz = LB(rsxz, D); O.szz = LB(rszz, D);
        O.hk[0] = bload(rsxz, vb + 8u + hd, pj + own);
        O.hk[1] = bload(rsyz, vb + 8u + hd, pj + own);
        O.hk[2] = bload(rszz, vb + 8u + hd, pj + own);
#pragma unroll
        for (int n = 0; n < NH; ++n) {
            const rsrc_t rc = oc[n] == 0 ? rsxx : (oc[n] == 1 ? rsxy : rsxz);
            O.ho[n] = bload(rc, vb + 8u, pj + ooff[n]);
        }
    };

    // j queues: sxy and syz at j-2..j+1, syy at j-1..j+2
    float sxy_m2 = LB(rsxy, -2), sxy_m1 = LB(rsxy, -1), sxy_0 = LB(rsxy, 0), sxy_p1 = LB(rsxy, 1);
    float syz_m2 = LB(rsyz, -2), syz_m1 = LB(rsyz, -1), syz_0 = LB(rsyz, 0), syz_p1 = LB(rsyz, 1);
    float syy_m1 = LB(rsyy, -1), syy_0 = LB(rsyy, 0), syy_p1 = LB(rsyy, 1), syy_p2 = LB(rsyy, 2);
    Ops cur, nxt{};
    load_ops(cur, 0);

    for (int j = jb; j <= je; ++j, vb += SJ) {
        const int p = (j - jb) & 1;
        // plane j+1's operands and the queue entries after it (sxy, syz at
        // j+2, syy at j+3 clamped to the last plane in memory), unconditional
        // as in k_s_lds
        load_ops(nxt, 1);
        load_cell(cur, 0);
        const uint32_t q3 = (uint32_t)(min(j + 3, g.ny + 2) - j + 2) * SJ + own;
        const float sxy_n = LB(rsxy, 2), syz_n = LB(rsyz, 2), syy_n = bload(rsyy, vb + 8u, q3);
        // stage plane j: halo words first, then the rows over them
        sh[p][2][y + 2][hcol] = cur.hk[0];
        sh[p][3][y + 2][hcol] = cur.hk[1];
        sh[p][4][y + 2][hcol] = cur.hk[2];
        sh[p][0][y + 2][lane + 2] = cur.sxx;
        sh[p][1][y + 2][lane + 2] = sxy_0;
        sh[p][2][y + 2][lane + 2] = cur.sxz;
        sh[p][3][y + 2][lane + 2] = syz_0;
        sh[p][4][y + 2][lane + 2] = cur.szz;
#pragma unroll
        for (int n = 0; n < NH; ++n) sh[p][oc[n]][orow[n] + 2][lane + 2] = cur.ho[n];
        __asm__ volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
        if (act) {
            const float(*XX)[68] = sh[p][0];
            const float(*XY)[68] = sh[p][1];
            const float(*XZ)[68] = sh[p][2];
            const float(*YZ)[68] = sh[p][3];
            const float(*ZZ)[68] = sh[p][4];
            const int c0 = lane + 2, r0_ = y + 2;
            float sxx_x = g.dtdx * D4F(b1, b2, XX[r0_ - 1][c0], cur.sxx, XX[r0_ + 1][c0], XX[r0_ + 2][c0]);
            float sxy_y = g.dtdy * D4B(b1, b2, sxy_m2, sxy_m1, sxy_0, sxy_p1);
            float sxz_z = g.dtdz * D4B(b1, b2, XZ[r0_][c0 - 2], XZ[r0_][c0 - 1], cur.sxz, XZ[r0_][c0 + 1]);
            float syy_y = g.dtdy * D4F(b1, b2, syy_m1, syy_0, syy_p1, syy_p2);
            float sxy_x = g.dtdx * D4B(b1, b2, XY[r0_ - 2][c0], XY[r0_ - 1][c0], sxy_0, XY[r0_ + 1][c0]);
            float syz_z = g.dtdz * D4B(b1, b2, YZ[r0_][c0 - 2], YZ[r0_][c0 - 1], syz_0, YZ[r0_][c0 + 1]);
            float szz_z = g.dtdz * D4F(b1, b2, ZZ[r0_][c0 - 1], cur.szz, ZZ[r0_][c0 + 1], ZZ[r0_][c0 + 2]);
            float sxz_x = g.dtdx * D4B(b1, b2, XZ[r0_ - 2][c0], XZ[r0_ - 1][c0], cur.sxz, XZ[r0_ + 1][c0]);
            float syz_y = g.dtdy * D4B(b1, b2, syz_m2, syz_m1, syz_0, syz_p1);
            if (CPML) {
                const int hy = slab_h(j, g.ny, g.fw, !g.free_surf);
                v_cpml(g, P, psi, j, i, k, hx, hy, hz, sxx_x, sxy_x, sxz_x, sxy_y, syy_y, syz_y, sxz_z, syz_z,
                       szz_z);
            }
            const float nvx = cur.vx0 + (sxx_x + sxy_y + sxz_z) / cur.rip;
            const float nvy = cur.vy0 + (syy_y + sxy_x + syz_z) / cur.rjp;
            const float nvz = cur.vz0 + (szz_z + sxz_x + syz_y) / cur.rkp;
            const uint32_t st = 2u * SJ + own;
            bstore_cp<CPN>(nvx, rvx, vb + 8u, st);
            bstore_cp<CPN>(nvy, rvy, vb + 8u, st);
            bstore_cp<CPN>(nvz, rvz, vb + 8u, st);
            if (SNAP) {
                bstore_cp<CPN>(nvx, rring, vb + 8u, st + N.soff);
                bstore_cp<CPN>(nvy, rring, vb + 8u, st + N.soff + g.fbytes);
                bstore_cp<CPN>(nvz, rring, vb + 8u, st + N.soff + 2 * g.fbytes);
            }
        }
        sxy_m2 = sxy_m1; sxy_m1 = sxy_0; sxy_0 = sxy_p1; sxy_p1 = sxy_n;
        syz_m2 = syz_m1; syz_m1 = syz_0; syz_0 = syz_p1; syz_p1 = syz_n;
        syy_m1 = syy_0; syy_0 = syy_p1; syy_p1 = syy_p2; syy_p2 = syy_n;
        cur = nxt;
    }
#undef LB
#undef LBN
}

// Operands of one plane j of the stress march (plus the j-queue values first
// needed at j: vx, vz at j+2, vy at j+1), and for the viscoelastic update
// the memory variables and relaxation ratios of the cell.
struct SPlane {
    float vx_xa, vx_xb, vx_xd, vx_zb, vx_zc, vx_zd, vy_xb, vy_xc, vy_xd, vy_zb, vy_zc, vy_zd, vz_xb, vz_xc, vz_xd,
        vz_za, vz_zb, vz_zd, uipjp, ujpkp, uipkp, pi, u, sxx0, syy0, szz0, sxy0, syz0, sxz0, vx_n, vz_n, vy_n;
    float r0[6], ts, tp, tsipjp, tsjpkp, tsipkp;
    float vx_h, vy_h, vz_h;  // k-row halos (krow wavefronts)
};
#define SLOAD(P, D)                                                                                            \
    do {                                                                                                       \
        P.vx_xa = LD(rvx, D, -2, 0); P.vx_xb = LD(rvx, D, -1, 0); P.vx_xd = LD(rvx, D, 1, 0);                    \
        P.vy_xb = LD(rvy, D, -1, 0); P.vy_xc = LD(rvy, D, 1, 0); P.vy_xd = LD(rvy, D, 2, 0);                     \
        P.vz_xb = LD(rvz, D, -1, 0); P.vz_xc = LD(rvz, D, 1, 0); P.vz_xd = LD(rvz, D, 2, 0);                     \
        if (krow) {                                                                                            \
            P.vx_h = LDH(rvx, D); P.vy_h = LDH(rvy, D); P.vz_h = LDH(rvz, D);                                  \
        } else {                                                                                               \
            P.vx_zb = LD(rvx, D, 0, -1); P.vx_zc = LD(rvx, D, 0, 1); P.vx_zd = LD(rvx, D, 0, 2);                 \
            P.vy_zb = LD(rvy, D, 0, -1); P.vy_zc = LD(rvy, D, 0, 1); P.vy_zd = LD(rvy, D, 0, 2);                 \
            P.vz_za = LD(rvz, D, 0, -2); P.vz_zb = LD(rvz, D, 0, -1); P.vz_zd = LD(rvz, D, 0, 1);                \
        }                                                                                                      \
        P.uipjp = LD(ruipjp, D, 0, 0); P.ujpkp = LD(rujpkp, D, 0, 0); P.uipkp = LD(ruipkp, D, 0, 0);             \
        P.pi = LD(rpi, D, 0, 0); P.u = LD(ru, D, 0, 0);                                                        \
        P.sxx0 = LDN(rsxx, D, 0, 0); P.syy0 = LDN(rsyy, D, 0, 0); P.szz0 = LDN(rszz, D, 0, 0);                   \
        P.sxy0 = LDN(rsxy, D, 0, 0); P.syz0 = LDN(rsyz, D, 0, 0); P.sxz0 = LDN(rsxz, D, 0, 0);                   \
        P.vx_n = LD(rvx, (D) + 2, 0, 0); P.vz_n = LD(rvz, (D) + 2, 0, 0); P.vy_n = LD(rvy, (D) + 1, 0, 0);       \
        if (VM) {                                                                                              \
            for (int q_ = 0; q_ < 6; ++q_) P.r0[q_] = LDN(rr[q_], D, 0, 0);                                    \
            P.ts = LD(rtaus, D, 0, 0); P.tp = LD(rtaup, D, 0, 0);                                              \
            P.tsipjp = LD(rtsipjp, D, 0, 0); P.tsjpkp = LD(rtsjpkp, D, 0, 0); P.tsipkp = LD(rtsipkp, D, 0, 0);  \
        }                                                                                                      \
    } while (0)

// VM: 0 elastic, 1 one relaxation mechanism (the reference's update_s),
// 2 V.L >= 1 mechanisms (stress_visco_n)
template <int TI, bool CPML, int VM, bool PIPE>
__global__ __launch_bounds__(64 * TI) void k_s_march(DevGeom g, MTiles T, SFields F, Prof P, Psi9 psi, Shots S,
                                                    ViscoF V)
{
    int k, i, jb, je, shot;
    Box R;
    bool krow;
    if (!decode_tile(T, TI, k, i, jb, je, shot, R, &krow)) return;
    const uint32_t khv = khalo_off();
    constexpr int CPN = CPML ? IFOS_CPOL_NT : 0;
    const uint32_t SJ = (uint32_t)g.sj * 4u, SI = (uint32_t)g.si * 4u;
    const float b1 = g.b1, b2 = g.b2;
    const float DT = g.dt;
    const int64_t so = (int64_t)shot * S.fstride;
    const rsrc_t rvx = make_rsrc(F.vx + so, g.fbytes), rvy = make_rsrc(F.vy + so, g.fbytes),
                 rvz = make_rsrc(F.vz + so, g.fbytes), rsxx = make_rsrc(F.sxx + so, g.fbytes),
                 rsyy = make_rsrc(F.syy + so, g.fbytes), rszz = make_rsrc(F.szz + so, g.fbytes),
                 rsxy = make_rsrc(F.sxy + so, g.fbytes), rsyz = make_rsrc(F.syz + so, g.fbytes),
                 rsxz = make_rsrc(F.sxz + so, g.fbytes), rpi = make_rsrc(F.pi, g.fbytes), ru = make_rsrc(F.u, g.fbytes),
                 ruipjp = make_rsrc(F.uipjp, g.fbytes), rujpkp = make_rsrc(F.ujpkp, g.fbytes),
                 ruipkp = make_rsrc(F.uipkp, g.fbytes);
    if (CPML) shift_psi(psi, S, shot);
    rsrc_t rr[6], rtaus, rtaup, rtsipjp, rtsjpkp, rtsipkp;
    if (VM) {
        const int nm = VM == 1 ? 1 : V.L;
#pragma unroll
        for (int n = 0; n < 6; ++n) rr[n] = make_rsrc(V.r[n] + so * nm, g.fbytes * (uint32_t)nm);
        rtaus = make_rsrc(V.taus, g.fbytes);
        rtaup = make_rsrc(V.taup, g.fbytes);
        rtsipjp = make_rsrc(V.tausipjp, g.fbytes);
        rtsjpkp = make_rsrc(V.tausjpkp, g.fbytes);
        rtsipkp = make_rsrc(V.tausipkp, g.fbytes);
    }
    uint32_t vo = (uint32_t)((int64_t)jb * g.sj + (int64_t)i * g.si + k + IFOS_KOFF - 2) * 4u;
    int hx = 0;
    if (CPML) hx = slab_h(i, g.nx, g.fw, 1);
    const int hz = CPML ? slab_h(k, g.nz, g.fw, 1) : 0;

    // j queues: vx, vz at j-1..j+1, vy at j-2..j
    float vx_m1 = LD(rvx, -1, 0, 0), vx_0 = LD(rvx, 0, 0, 0), vx_p1 = LD(rvx, 1, 0, 0);
    float vz_m1 = LD(rvz, -1, 0, 0), vz_0 = LD(rvz, 0, 0, 0), vz_p1 = LD(rvz, 1, 0, 0);
    float vy_m2 = LD(rvy, -2, 0, 0), vy_m1 = LD(rvy, -1, 0, 0), vy_0 = LD(rvy, 0, 0, 0);
    SPlane cur, nxt;
    SLOAD(cur, 0);

    for (int j = jb; j <= je; ++j, vo += SJ) {
        if (PIPE && j < je) SLOAD(nxt, 1);
        if (krow) {
            const KNb a = kshift(vx_0, cur.vx_h), b = kshift(vy_0, cur.vy_h), c = kshift(vz_0, cur.vz_h);
            cur.vx_zb = a.m1; cur.vx_zc = a.p1; cur.vx_zd = a.p2;
            cur.vy_zb = b.m1; cur.vy_zc = b.p1; cur.vy_zd = b.p2;
            cur.vz_za = c.m2; cur.vz_zb = c.m1; cur.vz_zd = c.p1;
        }
        float vxx = D4B(b1, b2, cur.vx_xa, cur.vx_xb, vx_0, cur.vx_xd);
        float vxy = D4F(b1, b2, vx_m1, vx_0, vx_p1, cur.vx_n);
        float vxz = D4F(b1, b2, cur.vx_zb, vx_0, cur.vx_zc, cur.vx_zd);
        float vyx = D4F(b1, b2, cur.vy_xb, vy_0, cur.vy_xc, cur.vy_xd);
        float vyy = D4B(b1, b2, vy_m2, vy_m1, vy_0, cur.vy_n);
        float vyz = D4F(b1, b2, cur.vy_zb, vy_0, cur.vy_zc, cur.vy_zd);
        float vzx = D4F(b1, b2, cur.vz_xb, vz_0, cur.vz_xc, cur.vz_xd);
        float vzy = D4F(b1, b2, vz_m1, vz_0, vz_p1, cur.vz_n);
        float vzz = D4B(b1, b2, cur.vz_za, cur.vz_zb, vz_0, cur.vz_zd);
        div_spacing(g, vxx, vxy, vxz, vyx, vyy, vyz, vzx, vzy, vzz);
        if (CPML) {
            const int hy = slab_h(j, g.ny, g.fw, !g.free_surf);
            s_cpml(g, P, psi, j, i, k, hx, hy, hz, vxx, vyx, vzx, vxy, vyy, vzy, vxz, vyz, vzz);
        }
        const float vxyyx = vxy + vyx, vyzzy = vyz + vzy, vxzzx = vxz + vzx;
        const float vxxyyzz = vxx + vyy + vzz;
        const float vyyzz = vyy + vzz, vxxzz = vxx + vzz, vxxyy = vxx + vyy;
        float sn[6], rn[6];
        if (!VM) {
            // update_s_ssg_elastic.c:117-138
            const float fipjp = cur.uipjp * DT, fjpkp = cur.ujpkp * DT, fipkp = cur.uipkp * DT;
            const float f = 2.0f * cur.u;
            sn[3] = cur.sxy0 + (fipjp * vxyyx);
            sn[4] = cur.syz0 + (fjpkp * vyzzy);
            sn[5] = cur.sxz0 + (fipkp * vxzzx);
            sn[0] = cur.sxx0 + DT * ((cur.pi * vxxyyzz) - (f * vyyzz));
            sn[1] = cur.syy0 + DT * ((cur.pi * vxxyyzz) - (f * vxxzz));
            sn[2] = cur.szz0 + DT * ((cur.pi * vxxyyzz) - (f * vxxyy));
        } else {
            sn[0] = cur.sxx0; sn[1] = cur.syy0; sn[2] = cur.szz0;
            sn[3] = cur.sxy0; sn[4] = cur.syz0; sn[5] = cur.sxz0;
#pragma unroll
            for (int n = 0; n < 6; ++n) rn[n] = cur.r0[n];
            if (VM == 1) {
                stress_visco(DT, V.eta[0], V.L, cur.pi, cur.u, cur.uipjp, cur.ujpkp, cur.uipkp, cur.ts, cur.tp,
                             cur.tsipjp, cur.tsjpkp, cur.tsipkp, vxyyx, vyzzy, vxzzx, vxxyyzz, vyyzz, vxxzz, vxxyy,
                             sn, rn);
            } else {
                const uint32_t own = 2 * SJ + 2 * SI;
                auto ldm = [&](int l, int n) { return bload_cp<CPN>(rr[n], vo + 8, own + (uint32_t)l * g.fbytes); };
                auto stm = [&](int l, int n, float v) {
                    bstore_cp<CPN>(v, rr[n], vo + 8, own + (uint32_t)l * g.fbytes);
                };
                stress_visco_n(DT, V, cur.pi, cur.u, cur.uipjp, cur.ujpkp, cur.uipkp, cur.ts, cur.tp, cur.tsipjp,
                               cur.tsjpkp, cur.tsipkp, vxyyx, vyzzy, vxzzx, vxxyyzz, vyyzz, vxxzz, vxxyy, sn, rn,
                               ldm, stm);
            }
        }
        vx_m1 = vx_0; vx_0 = vx_p1; vx_p1 = cur.vx_n;
        vz_m1 = vz_0; vz_0 = vz_p1; vz_p1 = cur.vz_n;
        vy_m2 = vy_m1; vy_m1 = vy_0; vy_0 = cur.vy_n;
        if (!PIPE && j < je) SLOAD(nxt, 1);
        STN(sn[3], rsxy);
        STN(sn[4], rsyz);
        STN(sn[5], rsxz);
        STN(sn[0], rsxx);
        STN(sn[1], rsyy);
        STN(sn[2], rszz);
        if (VM) {
#pragma unroll
            for (int n = 0; n < 6; ++n) STN(rn[n], rr[n]);
        }
        cur = nxt;
    }
}
#undef SLOAD

// ---------------------------------------------------------------------------
// Stress march with the in-plane neighbours staged in LDS.  Tiles of 64
// k-lanes (one row per wavefront, one shot per workgroup): at each plane j
// every wavefront writes its row of vx, vy, vz (the centre of its j-queues)
// to LDS, the halo words of its row (k0-2, k0-1, k0+64, k0+65), and its share
// of the rows just outside the workgroup (vx at i0-2, i0-1, i0+TI; vy and vz
// at i0-1, i0+TI, i0+TI+1); every i- and k-neighbour of the update is then
// read from LDS instead of through L1.  Software-pipelined one plane ahead:
// the halo words, the outside rows and the next j-queue entries of plane
// j+1 are loaded while plane j is computed, so the LDS writes of plane j wait
// only for loads issued an iteration earlier.  The own cell's operands (old
// stresses, moduli) are loaded at the top of their own iteration, where the
// staging, the barrier and the derivatives cover their latency: holding
// them a plane ahead too would cost 11 VGPRs (90 instead of 79 for the
// batched elastic march) and the sixth wave per SIMD that fits in 80.
// Two LDS buffers alternate by plane, so one barrier per plane suffices (a
// wavefront writes a buffer again only after the next barrier, which every
// wavefront reaches after its reads of that buffer).  Every lane loads, at a
// clamped cell (k <= nz+2, rows <= nx+2): an updating lane (k <= nz, i <= nx)
// reads only k-2..k+2 and rows i-2..i+2, none of them clamped, so the clamps
// only keep the loads of idle lanes in bounds; only lanes inside the box
// update and store.  The barrier waits for LDS only (s_waitcnt lgkmcnt(0)):
// the loads in flight for the next plane stay in flight across it.  Same
// arithmetic, in the same order, as k_s_march.
template <int TI, bool CPML, int VM>
__global__ __launch_bounds__(64 * TI) void k_s_lds(DevGeom g, MTiles T, SFields F, Prof P, Psi9 psi, Shots S,
                                                  ViscoF V)
{
    constexpr int RW = TI + 4;              // LDS rows: i0-2 .. i0+TI+1
    constexpr int NH = (9 + TI - 1) / TI;   // outside rows per wavefront (9 per workgroup)
    constexpr int CPN = CPML ? IFOS_CPOL_NT : 0;
    __shared__ float sh[2][3][RW][68];
    const int b = blockIdx.x;
    const int L = (b & 7) * T.per + (b >> 3);
    if (L >= T.total) return;  // whole workgroup
    const int gs = T.group * T.nsg;
    const int rem = L % gs;
    const int shot = rem / T.group;
    int t = (L / gs) * T.group + rem % T.group;
    if (t >= T.start[T.nb] || shot >= T.ns) return;  // whole workgroup
    int q = 0;
    while (q + 1 < T.nb && t >= T.start[q + 1]) ++q;
    t -= T.start[q];
    const Box R = T.b[q];
    const int tk = t % T.ntk[q];
    t /= T.ntk[q];
    const int ti = t % T.nti[q], tj = t / T.nti[q];
    const int lane = (int)threadIdx.x;
    const int y = __builtin_amdgcn_readfirstlane((int)threadIdx.y);  // one row per wavefront
    const int k0 = R.k0 + tk * 64, i0 = R.i0 + ti * TI;
    const int k = k0 + lane, i = i0 + y;
    const int jb = R.j0 + tj * T.jc, je = min(jb + T.jc - 1, R.j1);
    const bool act = k <= R.k1 && i <= R.i1;  // updates and stores
    const int kc = min(k, g.nz + 2);          // clamped load cell
    const int yc = min(i, g.nx + 2) - i0;     // clamped own row (wavefront-uniform)
    const uint32_t SJ = (uint32_t)g.sj * 4u, SI = (uint32_t)g.si * 4u;
    const float b1 = g.b1, b2 = g.b2, DT = g.dt;
    const int64_t so = (int64_t)shot * S.fstride;
    const rsrc_t rvx = make_rsrc(F.vx + so, g.fbytes), rvy = make_rsrc(F.vy + so, g.fbytes),
                 rvz = make_rsrc(F.vz + so, g.fbytes), rsxx = make_rsrc(F.sxx + so, g.fbytes),
                 rsyy = make_rsrc(F.syy + so, g.fbytes), rszz = make_rsrc(F.szz + so, g.fbytes),
                 rsxy = make_rsrc(F.sxy + so, g.fbytes), rsyz = make_rsrc(F.syz + so, g.fbytes),
                 rsxz = make_rsrc(F.sxz + so, g.fbytes), rpi = make_rsrc(F.pi, g.fbytes), ru = make_rsrc(F.u, g.fbytes),
                 ruipjp = make_rsrc(F.uipjp, g.fbytes), rujpkp = make_rsrc(F.ujpkp, g.fbytes),
                 ruipkp = make_rsrc(F.uipkp, g.fbytes);
    if (CPML) shift_psi(psi, S, shot);
    rsrc_t rr[6], rtaus, rtaup, rtsipjp, rtsjpkp, rtsipkp;
    if (VM) {
        const int nm = VM == 1 ? 1 : V.L;
#pragma unroll
        for (int n = 0; n < 6; ++n) rr[n] = make_rsrc(V.r[n] + so * nm, g.fbytes * (uint32_t)nm);
        rtaus = make_rsrc(V.taus, g.fbytes);
        rtaup = make_rsrc(V.taup, g.fbytes);
        rtsipjp = make_rsrc(V.tausipjp, g.fbytes);
        rtsjpkp = make_rsrc(V.tausjpkp, g.fbytes);
        rtsipkp = make_rsrc(V.tausipkp, g.fbytes);
    }
    // byte offset of (j-2, i0-2, kc-2) at plane j: operand (j+dj, i0+r, kc)
    // at vb + 8 + (dj+2)*SJ + (r+2)*SI
    uint32_t vb = (uint32_t)((int64_t)jb * g.sj + (int64_t)i0 * g.si + kc + IFOS_KOFF - 2) * 4u;
    const uint32_t own = (uint32_t)(yc + 2) * SI;
#define LB(r, dj) bload(r, vb + 8u, (uint32_t)((dj) + 2) * SJ + own)
#define LBN(r, dj) bload_cp<CPN>(r, vb + 8u, (uint32_t)((dj) + 2) * SJ + own)
    // halo words of the own row: lane l loads k0-2+l (lanes 0, 1), k0+64 and
    // k0+65 (lanes 62, 63); lanes 2-31 repeat lane 1's word and lanes 32-61
    // lane 62's (the load touches two lines).  Lanes 2-61 write theirs to
    // their own column, which the row write after it overwrites, so the 64
    // lanes write 64 different words.  hd is the word's byte offset from the
    // lane's own word (>= -4 (lane + 2): the sum vb + 8 + hd stays >= 0).
    const int kh = lane < 2 ? k0 - 2 + lane : lane < 32 ? k0 - 1 : lane < 63 ? k0 + 64 : k0 + 65;
    const int hcol = lane < 62 ? (lane < 2 ? lane : lane + 2) : lane + 4;
    const uint32_t hd = (uint32_t)((min(kh, g.nz + 2) - kc) * 4);
    // outside rows of this wavefront: e = min(y + n TI, 8) is (component e / 3,
    // row i0 + orow[e]); the load row is clamped to nx+2.  The 9 rows take
    // NH TI slots; the spare slots repeat row 8 (the same words, written with
    // the same values), so no wavefront branches around a load or a write.
    int oc[NH], orow[NH];
    uint32_t ooff[NH];
#pragma unroll
    for (int n = 0; n < NH; ++n) {
        const int e = min(y + n * TI, 8);
        const int c = e < 3 ? 0 : (e < 6 ? 1 : 2), m = e - 3 * c;
        const int r = c == 0 ? (m == 0 ? -2 : (m == 1 ? -1 : TI)) : (m == 0 ? -1 : (m == 1 ? TI : TI + 1));
        oc[n] = c;
        orow[n] = r;
        ooff[n] = (uint32_t)(min(i0 + r, g.nx + 2) - i0 + 2) * SI;
    }
    int hx = 0;
    if (CPML) hx = slab_h(i, g.nx, g.fw, 1);
    const int hz = CPML ? slab_h(k, g.nz, g.fw, 1) : 0;

    // operands of plane j + D: the own cell's, the halo words, the outside rows
    struct Ops {
        float s[6], pi, u, uipjp, ujpkp, uipkp, r[6], ts, tp, tsipjp, tsjpkp, tsipkp;
        float hk[3], ho[NH];
    };
    // the cell's own stresses and moduli, loaded at the top of their own
    // iteration
    auto load_cell = [&](Ops &O, int D) {
        O.s[0] = LBN(rsxx, D); O.s[1] = LBN(rsyy, D); O.s[2] = LBN(rszz, D);
        O.s[3] = LBN(rsxy, D); O.s[4] = LBN(rsyz, D); O.s[5] = LBN(rsxz, D);
        O.uipjp = LB(ruipjp, D); O.ujpkp = LB(rujpkp, D); O.uipkp = LB(ruipkp, D);
        O.pi = LB(rpi, D); O.u = LB(ru, D);
    };
    // a plane ahead: what plane j + D stages in LDS, and the viscoelastic
    // operands
    auto load_ops = [&](Ops &O, int D) {
        const uint32_t pj = (uint32_t)(D + 2) * SJ;
        if (VM) {
#pragma unroll
            for (int n = 0; n < 6; ++n) O.r[n] = LBN(rr[n], D);
            O.ts = LB(rtaus, D); O.tp = LB(rtaup, D);
            O.tsipjp = LB(rtsipjp, D); O.tsjpkp = LB(rtsjpkp, D); O.tsipkp = LB(rtsipkp, D);
        }
        O.hk[0] = bload(rvx, vb + 8u + hd, pj + own);
        O.hk[1] = bload(rvy, vb + 8u + hd, pj + own);
        O.hk[2] = bload(rvz, vb + 8u + hd, pj + own);
#pragma unroll
        for (int n = 0; n < NH; ++n) {
            const rsrc_t rc = oc[n] == 0 ? rvx : (oc[n] == 1 ? rvy : rvz);
            O.ho[n] = bload(rc, vb + 8u, pj + ooff[n]);
        }
    };

    // j queues of the own column: vx, vz at j-1..j+2, vy at j-2..j+1
    float vx_m1 = LB(rvx, -1), vx_0 = LB(rvx, 0), vx_p1 = LB(rvx, 1), vx_p2 = LB(rvx, 2);
    float vz_m1 = LB(rvz, -1), vz_0 = LB(rvz, 0), vz_p1 = LB(rvz, 1), vz_p2 = LB(rvz, 2);
    float vy_m2 = LB(rvy, -2), vy_m1 = LB(rvy, -1), vy_0 = LB(rvy, 0), vy_p1 = LB(rvy, 1);
    Ops cur, nxt{};
    load_ops(cur, 0);

    for (int j = jb; j <= je; ++j, vb += SJ) {
        const int p = (j - jb) & 1;
        // plane j+1's operands and the queue entries after it (vx, vz at j+3,
        // vy at j+2), loaded unconditionally (a branch here would make the
        // waits for this iteration's operands wait for these loads too): at
        // j = je they are unused, plane je+1 <= ny+1 exists and j+3 is
        // clamped to the last plane in memory, ny+2
        load_ops(nxt, 1);
        load_cell(cur, 0);
        const uint32_t q3 = (uint32_t)(min(j + 3, g.ny + 2) - j + 2) * SJ + own;
        const float vx_n = bload(rvx, vb + 8u, q3), vz_n = bload(rvz, vb + 8u, q3), vy_n = LB(rvy, 2);
        // stage plane j: halo words first, then the row over them
        sh[p][0][y + 2][hcol] = cur.hk[0];
        sh[p][1][y + 2][hcol] = cur.hk[1];
        sh[p][2][y + 2][hcol] = cur.hk[2];
        sh[p][0][y + 2][lane + 2] = vx_0;
        sh[p][1][y + 2][lane + 2] = vy_0;
        sh[p][2][y + 2][lane + 2] = vz_0;
#pragma unroll
        for (int n = 0; n < NH; ++n) sh[p][oc[n]][orow[n] + 2][lane + 2] = cur.ho[n];
        __asm__ volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
        if (act) {
            const float(*X)[68] = sh[p][0];
            const float(*Y)[68] = sh[p][1];
            const float(*Z)[68] = sh[p][2];
            const int c0 = lane + 2, r0_ = y + 2;
            float vxx = D4B(b1, b2, X[r0_ - 2][c0], X[r0_ - 1][c0], vx_0, X[r0_ + 1][c0]);
            float vxy = D4F(b1, b2, vx_m1, vx_0, vx_p1, vx_p2);
            float vxz = D4F(b1, b2, X[r0_][c0 - 1], vx_0, X[r0_][c0 + 1], X[r0_][c0 + 2]);
            float vyx = D4F(b1, b2, Y[r0_ - 1][c0], vy_0, Y[r0_ + 1][c0], Y[r0_ + 2][c0]);
            float vyy = D4B(b1, b2, vy_m2, vy_m1, vy_0, vy_p1);
            float vyz = D4F(b1, b2, Y[r0_][c0 - 1], vy_0, Y[r0_][c0 + 1], Y[r0_][c0 + 2]);
            float vzx = D4F(b1, b2, Z[r0_ - 1][c0], vz_0, Z[r0_ + 1][c0], Z[r0_ + 2][c0]);
            float vzy = D4F(b1, b2, vz_m1, vz_0, vz_p1, vz_p2);
            float vzz = D4B(b1, b2, Z[r0_][c0 - 2], Z[r0_][c0 - 1], vz_0, Z[r0_][c0 + 1]);
            div_spacing(g, vxx, vxy, vxz, vyx, vyy, vyz, vzx, vzy, vzz);
            if (CPML) {
                const int hy = slab_h(j, g.ny, g.fw, !g.free_surf);
                s_cpml(g, P, psi, j, i, k, hx, hy, hz, vxx, vyx, vzx, vxy, vyy, vzy, vxz, vyz, vzz);
            }
            const float vxyyx = vxy + vyx, vyzzy = vyz + vzy, vxzzx = vxz + vzx;
            const float vxxyyzz = vxx + vyy + vzz;
            const float vyyzz = vyy + vzz, vxxzz = vxx + vzz, vxxyy = vxx + vyy;
            float sn[6], rn[6];
            if (!VM) {
                // update_s_ssg_elastic.c:117-138
                const float fipjp = cur.uipjp * DT, fjpkp = cur.ujpkp * DT, fipkp = cur.uipkp * DT;
                const float f = 2.0f * cur.u;
                sn[3] = cur.s[3] + (fipjp * vxyyx);
                sn[4] = cur.s[4] + (fjpkp * vyzzy);
                sn[5] = cur.s[5] + (fipkp * vxzzx);
                sn[0] = cur.s[0] + DT * ((cur.pi * vxxyyzz) - (f * vyyzz));
                sn[1] = cur.s[1] + DT * ((cur.pi * vxxyyzz) - (f * vxxzz));
                sn[2] = cur.s[2] + DT * ((cur.pi * vxxyyzz) - (f * vxxyy));
            } else {
#pragma unroll
                for (int n = 0; n < 6; ++n) {
                    sn[n] = cur.s[n];
                    rn[n] = cur.r[n];
                }
                if (VM == 1) {
                    stress_visco(DT, V.eta[0], V.L, cur.pi, cur.u, cur.uipjp, cur.ujpkp, cur.uipkp, cur.ts, cur.tp,
                                 cur.tsipjp, cur.tsjpkp, cur.tsipkp, vxyyx, vyzzy, vxzzx, vxxyyzz, vyyzz, vxxzz,
                                 vxxyy, sn, rn);
                } else {
                    const uint32_t sto = 2u * SJ + own;
                    auto ldm = [&](int l, int n) {
                        return bload_cp<CPN>(rr[n], vb + 8u, sto + (uint32_t)l * g.fbytes);
                    };
                    auto stm = [&](int l, int n, float v) {
                        bstore_cp<CPN>(v, rr[n], vb + 8u, sto + (uint32_t)l * g.fbytes);
                    };
                    stress_visco_n(DT, V, cur.pi, cur.u, cur.uipjp, cur.ujpkp, cur.uipkp, cur.ts, cur.tp,
                                   cur.tsipjp, cur.tsjpkp, cur.tsipkp, vxyyx, vyzzy, vxzzx, vxxyyzz, vyyzz, vxxzz,
                                   vxxyy, sn, rn, ldm, stm);
                }
            }
            const uint32_t st = 2u * SJ + own;
            bstore_cp<CPN>(sn[3], rsxy, vb + 8u, st);
            bstore_cp<CPN>(sn[4], rsyz, vb + 8u, st);
            bstore_cp<CPN>(sn[5], rsxz, vb + 8u, st);
            bstore_cp<CPN>(sn[0], rsxx, vb + 8u, st);
            bstore_cp<CPN>(sn[1], rsyy, vb + 8u, st);
            bstore_cp<CPN>(sn[2], rszz, vb + 8u, st);
            if (VM) {
#pragma unroll
                for (int n = 0; n < 6; ++n) bstore_cp<CPN>(rn[n], rr[n], vb + 8u, st);
            }
        }
        vx_m1 = vx_0; vx_0 = vx_p1; vx_p1 = vx_p2; vx_p2 = vx_n;
        vz_m1 = vz_0; vz_0 = vz_p1; vz_p1 = vz_p2; vz_p2 = vz_n;
        vy_m2 = vy_m1; vy_m1 = vy_0; vy_0 = vy_p1; vy_p1 = vy_n;
        cur = nxt;
    }
#undef LB
#undef LBN
}
// ---------------------------------------------------------------------------
// gradient_F.c:79-139 for one frequency of one shot as a march over j: the
// y-derivatives of the 12 DFT fields come from 4-deep register queues and the
// three gradients are read, updated and written once per call.  Calls in
// frequency order (and shot order) repeat the reference's per-cell
// accumulation order exactly.  f[0..2] forward Re x y z, f[3..5] forward Im,
// f[6..8] adjoint Re, f[9..11] adjoint Im.
struct GradArgs {
    const float *f[12];
    float fd;  // 1 / (4 pi^2 f^2) as gradient_F.c:70-72 rounds it
    const float *rho, *pi, *u;
    float *g1, *g2, *g3;
    double *acc[3];  // GRAD_TD: lambda, mu and rho sums, one double per cell of the device layout
    double w;        // GRAD_TD: weight of the time sample (the time step)
};

// MODE GRAD_HESS: hess_F.c:28-192 instead, the same fields read as the
// forward set (f, fi) and one receiver-side set (b, bi), fd = 1 / (2 pi f) as
// hess_F.c:75-77 rounds it, the three Hessian diagonals in place of g1..g3.
// MODE GRAD_TD: one time sample of the time-domain gradient: f[0..2] the
// forward displacement, f[3..5] the forward velocity, f[6..8] the adjoint
// displacement, f[9..11] the adjoint velocity; the products of
// gradient_F.c:122-132 (displacement derivatives for lambda and mu, the
// velocities for rho, no frequency factor) are added, times w, to acc[0..2].
enum { GRAD_F = 0, GRAD_HESS = 1, GRAD_TD = 2 };

template <int TI, int MODE>
__global__ __launch_bounds__(64 * TI) void k_grad_march(DevGeom g, MTiles T, GradArgs A)
{
    int k, i, jb, je, shot;
    Box R;
    if (!decode_tile(T, TI, k, i, jb, je, shot, R)) return;
    const uint32_t SJ = (uint32_t)g.sj * 4u, SI = (uint32_t)g.si * 4u;
    const float b1 = g.b1, b2 = g.b2;
    rsrc_t rf[12];
#pragma unroll
    for (int n = 0; n < 12; ++n) rf[n] = make_rsrc(A.f[n], g.fbytes);
    const rsrc_t rrho = make_rsrc(A.rho, g.fbytes), rpi = make_rsrc(A.pi, g.fbytes), ru = make_rsrc(A.u, g.fbytes),
                 rg1 = make_rsrc(A.g1, g.fbytes), rg2 = make_rsrc(A.g2, g.fbytes), rg3 = make_rsrc(A.g3, g.fbytes);
    uint32_t vo = (uint32_t)((int64_t)jb * g.sj + (int64_t)i * g.si + k + IFOS_KOFF - 2) * 4u;
    // queues per triple t: x component at j-1..j+2, y at j-2..j+1, z at j-1..j+2
    float qx[4][4], qy[4][4], qz[4][4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        qx[t][0] = LD(rf[3 * t], -1, 0, 0); qx[t][1] = LD(rf[3 * t], 0, 0, 0); qx[t][2] = LD(rf[3 * t], 1, 0, 0);
        qy[t][0] = LD(rf[3 * t + 1], -2, 0, 0); qy[t][1] = LD(rf[3 * t + 1], -1, 0, 0);
        qy[t][2] = LD(rf[3 * t + 1], 0, 0, 0);
        qz[t][0] = LD(rf[3 * t + 2], -1, 0, 0); qz[t][1] = LD(rf[3 * t + 2], 0, 0, 0);
        qz[t][2] = LD(rf[3 * t + 2], 1, 0, 0);
    }
    for (int j = jb; j <= je; ++j, vo += SJ) {
        float d[4][9];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const rsrc_t fx = rf[3 * t], fy = rf[3 * t + 1], fz = rf[3 * t + 2];
            qx[t][3] = LD(fx, 2, 0, 0);
            qy[t][3] = LD(fy, 1, 0, 0);
            qz[t][3] = LD(fz, 2, 0, 0);
            if (MODE == GRAD_TD && (t & 1)) continue;  // velocities: cell values only
            const float x_c = qx[t][1], y_c = qy[t][2], z_c = qz[t][1];
            d[t][0] = b1 * (x_c - LD(fx, 0, -1, 0)) + b2 * (LD(fx, 0, 1, 0) - LD(fx, 0, -2, 0));
            d[t][1] = b1 * (qx[t][2] - x_c) + b2 * (qx[t][3] - qx[t][0]);
            d[t][2] = b1 * (LD(fx, 0, 0, 1) - x_c) + b2 * (LD(fx, 0, 0, 2) - LD(fx, 0, 0, -1));
            d[t][3] = b1 * (LD(fy, 0, 1, 0) - y_c) + b2 * (LD(fy, 0, 2, 0) - LD(fy, 0, -1, 0));
            d[t][4] = b1 * (y_c - qy[t][1]) + b2 * (qy[t][3] - qy[t][0]);
            d[t][5] = b1 * (LD(fy, 0, 0, 1) - y_c) + b2 * (LD(fy, 0, 0, 2) - LD(fy, 0, 0, -1));
            d[t][6] = b1 * (LD(fz, 0, 1, 0) - z_c) + b2 * (LD(fz, 0, 2, 0) - LD(fz, 0, -1, 0));
            d[t][7] = b1 * (qz[t][2] - z_c) + b2 * (qz[t][3] - qz[t][0]);
            d[t][8] = b1 * (z_c - LD(fz, 0, 0, -1)) + b2 * (LD(fz, 0, 0, 1) - LD(fz, 0, 0, -2));
            div_spacing(g, d[t][0], d[t][1], d[t][2], d[t][3], d[t][4], d[t][5], d[t][6], d[t][7], d[t][8]);
        }
        // d[t] = xx xy xz yx yy yz zx zy zz of f (t 0), fi (1), b (2), bi (3)
        const float *f = d[0], *fi = d[1], *bb = d[2], *bi = d[3];
        if (MODE == GRAD_TD) {
            float gl = (f[0] + f[4] + f[8]) * (bb[0] + bb[4] + bb[8]);
            gl = -gl;
            float gm = 2.0f * f[0] * bb[0] + 2.0f * f[4] * bb[4] + 2.0f * f[8] * bb[8] + (f[1] + f[3]) * (bb[1] + bb[3]) +
                       (f[2] + f[6]) * (bb[2] + bb[6]) + (f[5] + f[7]) * (bb[5] + bb[7]);
            gm = -gm;
            const float gr = qx[1][1] * qx[3][1] + qy[1][2] * qy[3][2] + qz[1][1] * qz[3][1];
            const int64_t c = at(g, j, i, k);
            A.acc[0][c] += (double)gl * A.w;
            A.acc[1][c] += (double)gm * A.w;
            A.acc[2][c] += (double)gr * A.w;
        } else if (MODE == GRAD_HESS) {
            // hess_F.c:135-182, values of the cell: qx[t][1], qy[t][2], qz[t][1]
            const float fdiv = f[0] + f[4] + f[8], fidiv = fi[0] + fi[4] + fi[8];
            const float bdiv = bb[0] + bb[4] + bb[8], bidiv = bi[0] + bi[4] + bi[8];
            float relam = fidiv * bdiv + fdiv * bidiv;
            relam = A.fd * relam;
            float imlam = fidiv * bidiv - fdiv * bdiv;
            imlam = A.fd * imlam;
            float remu = 2.0f * fi[0] * bb[0] + 2.0f * fi[4] * bb[4] + 2.0f * fi[8] * bb[8] + 2.0f * f[0] * bi[0] +
                         2.0f * f[4] * bi[4] + 2.0f * f[8] * bi[8] + (fi[1] + fi[3]) * (bb[1] + bb[3]) +
                         (f[1] + f[3]) * (bi[1] + bi[3]) + (fi[2] + fi[6]) * (bb[2] + bb[6]) +
                         (f[2] + f[6]) * (bi[2] + bi[6]) + (fi[5] + fi[7]) * (bb[5] + bb[7]) +
                         (f[5] + f[7]) * (bi[5] + bi[7]);
            remu = A.fd * remu;
            float immu = 2.0f * fi[0] * bi[0] + 2.0f * fi[4] * bi[4] + 2.0f * fi[8] * bi[8] - 2.0f * f[0] * bb[0] -
                         2.0f * f[4] * bb[4] - 2.0f * f[8] * bb[8] + (fi[1] + fi[3]) * (bi[1] + bi[3]) -
                         (f[1] + f[3]) * (bb[1] + bb[3]) + (fi[2] + fi[6]) * (bi[2] + bi[6]) -
                         (f[2] + f[6]) * (bb[2] + bb[6]) + (fi[5] + fi[7]) * (bi[5] + bi[7]) -
                         (f[5] + f[7]) * (bb[5] + bb[7]);
            immu = A.fd * immu;
            const float fx = qx[0][1], fy = qy[0][2], fz = qz[0][1], fix = qx[1][1], fiy = qy[1][2], fiz = qz[1][1];
            const float bx = qx[2][1], by = qy[2][2], bz = qz[2][1], bix = qx[3][1], biy = qy[3][2], biz = qz[3][1];
            const float rerho = (fx * bx + fy * by + fz * bz) - (fix * bix + fiy * biy + fiz * biz);
            const float imrho = (fx * bix + fy * biy + fz * biz) + (fix * bx + fiy * by + fiz * bz);
            const float revs = -2.0f * relam + remu, imvs = -2.0f * imlam + immu;
            const float r = LD(rrho, 0, 0, 0), p = LD(rpi, 0, 0, 0), m = LD(ru, 0, 0, 0);
            const float rerho1 = rerho + relam * (p - 2.0f * m) / r + remu * m / r;
            const float imrho1 = imrho + imlam * (p - 2.0f * m) / r + immu * m / r;
            float h1 = LD(rg1, 0, 0, 0), h2 = LD(rg2, 0, 0, 0), h3 = LD(rg3, 0, 0, 0);
            h1 += (relam * relam + imlam * imlam) * r * p * 4.0f;
            h2 += (revs * revs + imvs * imvs) * 4.0f * r * m;
            h3 += rerho1 * rerho1 + imrho1 * imrho1;
            ST(h1, rg1);
            ST(h2, rg2);
            ST(h3, rg3);
        } else {
        float gl = (f[0] + f[4] + f[8]) * (bb[0] + bb[4] + bb[8]) + (fi[0] + fi[4] + fi[8]) * (bi[0] + bi[4] + bi[8]);
        gl = -gl * A.fd;
        float gm = 2.0f * f[0] * bb[0] + 2.0f * f[4] * bb[4] + 2.0f * f[8] * bb[8] + 2.0f * fi[0] * bi[0] +
                   2.0f * fi[4] * bi[4] + 2.0f * fi[8] * bi[8] + (f[1] + f[3]) * (bb[1] + bb[3]) +
                   (fi[1] + fi[3]) * (bi[1] + bi[3]) + (f[2] + f[6]) * (bb[2] + bb[6]) +
                   (fi[2] + fi[6]) * (bi[2] + bi[6]) + (f[5] + f[7]) * (bb[5] + bb[7]) +
                   (fi[5] + fi[7]) * (bi[5] + bi[7]);
        gm = -gm * A.fd;
        const float gr = (qx[0][1] * qx[2][1] + qy[0][2] * qy[2][2] + qz[0][1] * qz[2][1]) +
                         (qx[1][1] * qx[3][1] + qy[1][2] * qy[3][2] + qz[1][1] * qz[3][1]);
        const float r = LD(rrho, 0, 0, 0), p = LD(rpi, 0, 0, 0), m = LD(ru, 0, 0, 0);
        float a1 = LD(rg1, 0, 0, 0), a2 = LD(rg2, 0, 0, 0), a3 = LD(rg3, 0, 0, 0);
        const double srp = sqrt((double)(r * p)), sru = sqrt((double)(r * m));
        a1 = (float)(a1 + srp * 2 * gl);
        a2 = (float)(a2 + (-4 * sru * gl + 2 * sru * gm));
        a3 += gr + m / r * gm + (p - 2.0f * m) / r * gl;
        ST(a1, rg1);
        ST(a2, rg2);
        ST(a3, rg3);
        }
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                qx[t][q] = qx[t][q + 1];
                qy[t][q] = qy[t][q + 1];
                qz[t][q] = qz[t][q + 1];
            }
    }
}

#undef LD
#undef ST
#undef LDH
#undef LDN
#undef STN

// ---------------------------------------------------------------------------
// slab kernels: one thread per CPML cell, k fastest within each box
__device__ inline bool slab_cell(const BoxList &L, int64_t t, int &j, int &i, int &k)
{
    if (t >= L.start[L.n]) return false;
    int b = 0;
    while (t >= L.start[b + 1]) ++b;
    const Box &B = L.b[b];
    int64_t r = t - L.start[b];
    const int nk = B.k1 - B.k0 + 1, ni = B.i1 - B.i0 + 1;
    k = B.k0 + (int)(r % nk);
    r /= nk;
    i = B.i0 + (int)(r % ni);
    j = B.j0 + (int)(r / ni);
    return true;
}

__global__ __launch_bounds__(256) void k_v_slabs(DevGeom g, BoxList L, VFields F, Prof P, Psi9 psi)
{
    int j, i, k;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (!slab_cell(L, t, j, i, k)) return;
    const int64_t c = at(g, j, i, k), sj = g.sj, si = g.si;
    const float b1 = g.b1, b2 = g.b2;
    float sxx_x = g.dtdx * D4F(b1, b2, F.sxx[c - si], F.sxx[c], F.sxx[c + si], F.sxx[c + 2 * si]);
    float sxy_x = g.dtdx * D4B(b1, b2, F.sxy[c - 2 * si], F.sxy[c - si], F.sxy[c], F.sxy[c + si]);
    float sxz_x = g.dtdx * D4B(b1, b2, F.sxz[c - 2 * si], F.sxz[c - si], F.sxz[c], F.sxz[c + si]);
    float sxy_y = g.dtdy * D4B(b1, b2, F.sxy[c - 2 * sj], F.sxy[c - sj], F.sxy[c], F.sxy[c + sj]);
    float syy_y = g.dtdy * D4F(b1, b2, F.syy[c - sj], F.syy[c], F.syy[c + sj], F.syy[c + 2 * sj]);
    float syz_y = g.dtdy * D4B(b1, b2, F.syz[c - 2 * sj], F.syz[c - sj], F.syz[c], F.syz[c + sj]);
    float sxz_z = g.dtdz * D4B(b1, b2, F.sxz[c - 2], F.sxz[c - 1], F.sxz[c], F.sxz[c + 1]);
    float syz_z = g.dtdz * D4B(b1, b2, F.syz[c - 2], F.syz[c - 1], F.syz[c], F.syz[c + 1]);
    float szz_z = g.dtdz * D4F(b1, b2, F.szz[c - 1], F.szz[c], F.szz[c + 1], F.szz[c + 2]);
    const int hx = slab_h(i, g.nx, g.fw, 1);
    const int hy = slab_h(j, g.ny, g.fw, !g.free_surf);
    const int hz = slab_h(k, g.nz, g.fw, 1);
    v_cpml(g, P, psi, j, i, k, hx, hy, hz, sxx_x, sxy_x, sxz_x, sxy_y, syy_y, syz_y, sxz_z, syz_z, szz_z);
    F.vx[c] += (sxx_x + sxy_y + sxz_z) / F.rip[c];
    F.vy[c] += (syy_y + sxy_x + syz_z) / F.rjp[c];
    F.vz[c] += (szz_z + sxz_x + syz_y) / F.rkp[c];
}

template <int VM>
__global__ __launch_bounds__(256) void k_s_slabs(DevGeom g, BoxList L, SFields F, Prof P, Psi9 psi, ViscoF V)
{
    int j, i, k;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (!slab_cell(L, t, j, i, k)) return;
    const int64_t c = at(g, j, i, k), sj = g.sj, si = g.si;
    const float b1 = g.b1, b2 = g.b2;

    float vxx = D4B(b1, b2, F.vx[c - 2 * si], F.vx[c - si], F.vx[c], F.vx[c + si]);
    float vxy = D4F(b1, b2, F.vx[c - sj], F.vx[c], F.vx[c + sj], F.vx[c + 2 * sj]);
    float vxz = D4F(b1, b2, F.vx[c - 1], F.vx[c], F.vx[c + 1], F.vx[c + 2]);
    float vyx = D4F(b1, b2, F.vy[c - si], F.vy[c], F.vy[c + si], F.vy[c + 2 * si]);
    float vyy = D4B(b1, b2, F.vy[c - 2 * sj], F.vy[c - sj], F.vy[c], F.vy[c + sj]);
    float vyz = D4F(b1, b2, F.vy[c - 1], F.vy[c], F.vy[c + 1], F.vy[c + 2]);
    float vzx = D4F(b1, b2, F.vz[c - si], F.vz[c], F.vz[c + si], F.vz[c + 2 * si]);
    float vzy = D4F(b1, b2, F.vz[c - sj], F.vz[c], F.vz[c + sj], F.vz[c + 2 * sj]);
    float vzz = D4B(b1, b2, F.vz[c - 2], F.vz[c - 1], F.vz[c], F.vz[c + 1]);
    div_spacing(g, vxx, vxy, vxz, vyx, vyy, vyz, vzx, vzy, vzz);
    const int hx = slab_h(i, g.nx, g.fw, 1);
    const int hy = slab_h(j, g.ny, g.fw, !g.free_surf);
    const int hz = slab_h(k, g.nz, g.fw, 1);
    s_cpml(g, P, psi, j, i, k, hx, hy, hz, vxx, vyx, vzx, vxy, vyy, vzy, vxz, vyz, vzz);
    if (!VM) {
        s_apply(g, F, c, vxx, vxy, vxz, vyx, vyy, vyz, vzx, vzy, vzz);
        return;
    }
    float sn[6] = {F.sxx[c], F.syy[c], F.szz[c], F.sxy[c], F.syz[c], F.sxz[c]};
    float rn[6];
#pragma unroll
    for (int n = 0; n < 6; ++n) rn[n] = V.r[n][c];
    if (VM == 1) {
        stress_visco(g.dt, V.eta[0], V.L, F.pi[c], F.u[c], F.uipjp[c], F.ujpkp[c], F.uipkp[c], V.taus[c],
                     V.taup[c], V.tausipjp[c], V.tausjpkp[c], V.tausipkp[c], vxy + vyx, vyz + vzy, vxz + vzx,
                     vxx + vyy + vzz, vyy + vzz, vxx + vzz, vxx + vyy, sn, rn);
    } else {
        auto ldm = [&](int l, int n) { return V.r[n][c + l * V.lstride]; };
        auto stm = [&](int l, int n, float v) { V.r[n][c + l * V.lstride] = v; };
        stress_visco_n(g.dt, V, F.pi[c], F.u[c], F.uipjp[c], F.ujpkp[c], F.uipkp[c], V.taus[c], V.taup[c],
                       V.tausipjp[c], V.tausjpkp[c], V.tausipkp[c], vxy + vyx, vyz + vzy, vxz + vzx, vxx + vyy + vzz,
                       vyy + vzz, vxx + vzz, vxx + vyy, sn, rn, ldm, stm);
    }
    F.sxx[c] = sn[0];
    F.syy[c] = sn[1];
    F.szz[c] = sn[2];
    F.sxy[c] = sn[3];
    F.syz[c] = sn[4];
    F.sxz[c] = sn[5];
#pragma unroll
    for (int n = 0; n < 6; ++n) V.r[n][c] = rn[n];
}

// ---------------------------------------------------------------------------
// host launchers

// the CPML slabs as boxes (update_v_ssg_CPML.c:70-375 loop bounds); the
// y-low slab is absent with a free surface
static int slab_boxes(const ifos_geom *g, Box *out)
{
    const int fw = g->fw, nx = g->nx, ny = g->ny, nz = g->nz;
    const int xb0 = fw + 1, xb1 = nx - fw, yb0 = g->free_surf ? 1 : fw + 1, yb1 = ny - fw;
    const Box all[6] = {
        {1, fw, 1, ny, 1, nz},                 // x low   (update_v_ssg_CPML.c:70)
        {xb1 + 1, nx, 1, ny, 1, nz},           // x high  (:142)
        {xb0, xb1, 1, fw, 1, nz},              // y low   (:215, no free surface)
        {xb0, xb1, yb1 + 1, ny, 1, nz},        // y high  (:267)
        {xb0, xb1, yb0, yb1, 1, fw},           // z low   (:320)
        {xb0, xb1, yb0, yb1, nz - fw + 1, nz}, // z high  (:352)
    };
    int n = 0;
    for (int b = 0; b < 6; ++b) {
        if (b == 2 && g->free_surf) continue;
        const Box &B = all[b];
        if (B.i1 < B.i0 || B.j1 < B.j0 || B.k1 < B.k0) continue;
        out[n++] = B;
    }
    return n;
}

// the slabs as a cell list for the one-thread-per-cell kernels
static BoxList slab_list(const ifos_geom *g)
{
    BoxList L;
    Box b[6];
    L.n = slab_boxes(g, b);
    L.start[0] = 0;
    for (int q = 0; q < L.n; ++q) {
        L.b[q] = b[q];
        L.start[q + 1] = L.start[q] + (int64_t)(b[q].i1 - b[q].i0 + 1) * (b[q].j1 - b[q].j0 + 1) *
                                          (b[q].k1 - b[q].k0 + 1);
    }
    for (int q = L.n + 1; q < 7; ++q) L.start[q] = L.start[L.n];
    return L;
}

static Box region_box(const ifos_geom *g, int region)
{
    if (region == IFOS_REGION_INTERIOR) {
        Box B = {g->fw + 1, g->nx - g->fw, g->free_surf ? 1 : g->fw + 1, g->ny - g->fw, g->fw + 1,
                 g->nz - g->fw};
        return B;
    }
    Box B = {1, g->nx, 1, g->ny, 1, g->nz};
    return B;
}

// tiles of boxes for TI i-rows per block; jc <= 0 picks the j chunk so a
// launch has >= 4096 workgroups (16 per CU) when the boxes allow it
static MTiles make_tiles(const Box *boxes, int nb, int ti, int jc, int ns, int group, int kb = 1, int sb = 1)
{
    MTiles T = {};
    T.nb = 0;
    T.ns = ns;
    // shots side by side in a workgroup share its model loads
    int sbq = sb < 1 ? 1 : sb;
    while (sbq > 1 && (sbq > ns || ti % sbq)) sbq /= 2;
    T.sb = sbq;
    T.nsg = (ns + sbq - 1) / sbq;
    T.start[0] = 0;
    for (int q = 0; q < nb; ++q) {
        const Box &B = boxes[q];
        if (B.i1 < B.i0 || B.j1 < B.j0 || B.k1 < B.k0) continue;
        const int nk = B.k1 - B.k0 + 1, ni = B.i1 - B.i0 + 1;
        const int kw = nk >= 64 ? 64 : nk, rpw = 64 / kw;
        const int nwk = (nk + kw - 1) / kw;  // wavefront columns along k
        int kbq = kb < 1 ? 1 : kb;
        while (kbq > 1 && (kbq > nwk || (ti / sbq) % kbq)) kbq /= 2;
        const int rows = (ti / (kbq * sbq)) * rpw;  // i-rows of a workgroup
        T.b[T.nb] = B;
        T.kw[T.nb] = kw;
        T.rpw[T.nb] = rpw;
        T.kb[T.nb] = kbq;
        T.ntk[T.nb] = (nwk + kbq - 1) / kbq;
        T.nti[T.nb] = (ni + rows - 1) / rows;
        ++T.nb;
    }
    auto count = [&](int c) {
        int64_t tot = 0;
        for (int q = 0; q < T.nb; ++q)
            tot += (int64_t)T.ntk[q] * T.nti[q] * ((T.b[q].j1 - T.b[q].j0 + 1 + c - 1) / c);
        return tot;
    };
    if (jc <= 0) {
        jc = 64;
        while (jc > 8 && count(jc) * T.nsg < 4096) jc /= 2;
    }
    T.jc = jc;
    for (int q = 0; q < T.nb; ++q)
        T.start[q + 1] = T.start[q] + T.ntk[q] * T.nti[q] * ((T.b[q].j1 - T.b[q].j0 + 1 + jc - 1) / jc);
    for (int q = T.nb + 1; q < 7; ++q) T.start[q] = T.start[T.nb];
    const int tiles = T.start[T.nb];
    T.group = group <= 0 ? tiles : std::min(group, tiles);
    if (T.group < 1) T.group = 1;
    T.total = ((tiles + T.group - 1) / T.group) * T.group * T.nsg;
    T.per = (T.total + 7) / 8;
    return T;
}

// default tile shape of the march kernels (i-rows per block, j chunk,
// shot grouping, ..., LDS staging where every box row holds 64 k-lanes);
// the batched path's autotuner overrides it per kernel.  On the 256^3
// interior of a 276^3 grid the LDS marches with 32-plane chunks run 4-8 %
// faster than the register marches over the whole j extent.
static MarchCfg g_default_cfg = {4, 32, 0, 0, 0, 1, 1, 1};

// the shape checks of ifos_batch_set_tiles and ifos_set_plain_tiles
int check_tile_cfg(const int *c, const char *who)
{
    if (c[0] != 1 && c[0] != 2 && c[0] != 4 && c[0] != 8) {
        set_error("%s: TI must be 1, 2, 4 or 8", who);
        return IFOS_EINVAL;
    }
    if (c[5] < 1 || c[6] < 1 || c[5] * c[6] > c[0] || c[0] % (c[5] * c[6])) {
        set_error("%s: kb=%d times sb=%d must divide TI=%d", who, c[5], c[6], c[0]);
        return IFOS_EINVAL;
    }
    return IFOS_OK;
}

template <bool CPML, bool SNAP, bool PIPE>
static void launch_v_ti(int ti, dim3 grid, const MTiles &T, const DevGeom &G, const VFields &F, const Prof &P,
                        const Psi9 &psi, const Shots &S, const Snap &N, hipStream_t s)
{
    switch (ti) {
    case 1: hipLaunchKernelGGL((k_v_march<1, CPML, SNAP, PIPE>), grid, dim3(64, 1), 0, s, G, T, F, P, psi, S, N); break;
    case 2: hipLaunchKernelGGL((k_v_march<2, CPML, SNAP, PIPE>), grid, dim3(64, 2), 0, s, G, T, F, P, psi, S, N); break;
    case 4: hipLaunchKernelGGL((k_v_march<4, CPML, SNAP, PIPE>), grid, dim3(64, 4), 0, s, G, T, F, P, psi, S, N); break;
    default: hipLaunchKernelGGL((k_v_march<8, CPML, SNAP, PIPE>), grid, dim3(64, 8), 0, s, G, T, F, P, psi, S, N); break;
    }
}

// the LDS-staged marches need 64-lane rows (every box at least 64 cells deep
// in k), one k column and one shot per workgroup, and TI >= 2
static bool lds_ok(const MarchCfg &c, const MTiles &T)
{
    if (!c.lds || c.ti < 2 || T.sb != 1) return false;
    for (int q = 0; q < T.nb; ++q)
        if (T.kw[q] != 64 || T.kb[q] != 1) return false;
    return true;
}

template <bool CPML, bool SNAP>
static void launch_v_lds_ti(int ti, dim3 grid, const MTiles &T, const DevGeom &G, const VFields &F, const Prof &P,
                            const Psi9 &psi, const Shots &S, const Snap &N, hipStream_t s)
{
    switch (ti) {
    case 2: hipLaunchKernelGGL((k_v_lds<2, CPML, SNAP>), grid, dim3(64, 2), 0, s, G, T, F, P, psi, S, N); break;
    case 4: hipLaunchKernelGGL((k_v_lds<4, CPML, SNAP>), grid, dim3(64, 4), 0, s, G, T, F, P, psi, S, N); break;
    default: hipLaunchKernelGGL((k_v_lds<8, CPML, SNAP>), grid, dim3(64, 8), 0, s, G, T, F, P, psi, S, N); break;
    }
}

template <bool CPML, bool SNAP>
static int launch_v(const MarchCfg &c, const MTiles &T, const DevGeom &G, const VFields &F, const Prof &P,
                    const Psi9 &psi, const Shots &S, const Snap &N, hipStream_t s)
{
    if (T.total == 0) return IFOS_OK;
    if (c.ti != 1 && c.ti != 2 && c.ti != 4 && c.ti != 8) {
        set_error("march: TI=%d not in {1,2,4,8}", c.ti);
        return IFOS_EINVAL;
    }
    const dim3 grid((unsigned)(T.per * 8));
    if (lds_ok(c, T)) {
        launch_v_lds_ti<CPML, SNAP>(c.ti, grid, T, G, F, P, psi, S, N, s);
        return hip_check(hipGetLastError(), "k_v_lds");
    }
    if (c.pipe) launch_v_ti<CPML, SNAP, true>(c.ti, grid, T, G, F, P, psi, S, N, s);
    else launch_v_ti<CPML, SNAP, false>(c.ti, grid, T, G, F, P, psi, S, N, s);
    return hip_check(hipGetLastError(), "k_v_march");
}

template <bool CPML, int VM, bool PIPE>
static void launch_s_ti(int ti, dim3 grid, const MTiles &T, const DevGeom &G, const SFields &F, const Prof &P,
                        const Psi9 &psi, const Shots &S, const ViscoF &V, hipStream_t s)
{
    switch (ti) {
    case 1: hipLaunchKernelGGL((k_s_march<1, CPML, VM, PIPE>), grid, dim3(64, 1), 0, s, G, T, F, P, psi, S, V); break;
    case 2: hipLaunchKernelGGL((k_s_march<2, CPML, VM, PIPE>), grid, dim3(64, 2), 0, s, G, T, F, P, psi, S, V); break;
    case 4: hipLaunchKernelGGL((k_s_march<4, CPML, VM, PIPE>), grid, dim3(64, 4), 0, s, G, T, F, P, psi, S, V); break;
    default: hipLaunchKernelGGL((k_s_march<8, CPML, VM, PIPE>), grid, dim3(64, 8), 0, s, G, T, F, P, psi, S, V); break;
    }
}

template <bool CPML, int VM>
static void launch_s_lds_ti(int ti, dim3 grid, const MTiles &T, const DevGeom &G, const SFields &F, const Prof &P,
                            const Psi9 &psi, const Shots &S, const ViscoF &V, hipStream_t s)
{
    switch (ti) {
    case 2: hipLaunchKernelGGL((k_s_lds<2, CPML, VM>), grid, dim3(64, 2), 0, s, G, T, F, P, psi, S, V); break;
    case 4: hipLaunchKernelGGL((k_s_lds<4, CPML, VM>), grid, dim3(64, 4), 0, s, G, T, F, P, psi, S, V); break;
    default: hipLaunchKernelGGL((k_s_lds<8, CPML, VM>), grid, dim3(64, 8), 0, s, G, T, F, P, psi, S, V); break;
    }
}

template <bool CPML, int VM = 0>
static int launch_s(const MarchCfg &c, const MTiles &T, const DevGeom &G, const SFields &F, const Prof &P,
                    const Psi9 &psi, const Shots &S, hipStream_t s, const ViscoF &V = ViscoF{})
{
    if (T.total == 0) return IFOS_OK;
    if (c.ti != 1 && c.ti != 2 && c.ti != 4 && c.ti != 8) {
        set_error("march: TI=%d not in {1,2,4,8}", c.ti);
        return IFOS_EINVAL;
    }
    const dim3 grid((unsigned)(T.per * 8));
    if (lds_ok(c, T)) {
        launch_s_lds_ti<CPML, VM>(c.ti, grid, T, G, F, P, psi, S, V, s);
        return hip_check(hipGetLastError(), "k_s_lds");
    }
    if (c.pipe) launch_s_ti<CPML, VM, true>(c.ti, grid, T, G, F, P, psi, S, V, s);
    else launch_s_ti<CPML, VM, false>(c.ti, grid, T, G, F, P, psi, S, V, s);
    return hip_check(hipGetLastError(), "k_s_march");
}

static VFields vfields(const ifos_fields *f, const ifos_model *m)
{
    VFields F = {f->vx, f->vy, f->vz, f->sxx, f->syy, f->szz, f->sxy, f->syz, f->sxz, m->rip, m->rjp, m->rkp};
    return F;
}
static SFields sfields(const ifos_fields *f, const ifos_model *m)
{
    SFields F = {f->vx, f->vy, f->vz, f->sxx, f->syy, f->szz, f->sxy, f->syz, f->sxz,
                 m->pi, m->u, m->uipjp, m->ujpkp, m->uipkp};
    return F;
}
static ViscoF viscof(const ifos_geom *g, const ifos_visco *v)
{
    ViscoF V = {};
    for (int n = 0; n < 6; ++n) V.r[n] = v->r[n];
    V.taus = v->taus;
    V.taup = v->taup;
    V.tausipjp = v->tausipjp;
    V.tausjpkp = v->tausjpkp;
    V.tausipkp = v->tausipkp;
    V.L = v->L;
    V.eta[0] = v->eta;
    for (int l = 1; l < IFOS_MAX_L; ++l) V.eta[l] = v->eta_more[l - 1];
    V.lstride = (int64_t)ifos_field_elems(g);
    return V;
}

static Psi9 psi9(const ifos_psi *p)
{
    Psi9 q = {};
    if (p)
        for (int n = 0; n < 9; ++n) q.p[n] = p->p[n];
    return q;
}

// one shot, one region: INTERIOR = the interior box, CPML = the six slabs,
// ALL = [1..N]^3 in one launch
static int update_v_region(const ifos_geom *g, int region, const Box *box, const ifos_fields *f,
                           const ifos_model *m, const float *d_prof, const ifos_psi *psi_s, hipStream_t s)
{
    DevGeom G = make_dev_geom(g);
    const VFields F = vfields(f, m);
    const Prof P = {d_prof, 2 * g->fw + 1};
    const Psi9 psi = psi9(psi_s);
    const Shots one = {1, 0, {0, 0, 0}};
    const Snap none = {nullptr, 0, 0, 0};
    const MarchCfg &c = g_default_cfg;
    if (region == IFOS_REGION_CPML && !box) {
        const BoxList L = slab_list(g);
        const int64_t n = L.start[L.n];
        if (n == 0) return IFOS_OK;
        ProfScope ps_(IFOS_K_V_CPML, s);
        hipLaunchKernelGGL(k_v_slabs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, G, L, F, P, psi);
        return hip_check(hipGetLastError(), "k_v_slabs");
    }
    Box boxes[1];
    boxes[0] = box ? *box : region_box(g, region);
    const MTiles T = make_tiles(boxes, 1, c.ti, c.jc, 1, 0);
    const int cat = region == IFOS_REGION_ALL ? IFOS_K_V_FUSED : IFOS_K_V_INTERIOR;
    ProfScope ps_(cat, s);
    if (region == IFOS_REGION_INTERIOR) return launch_v<false, false>(c, T, G, F, P, psi, one, none, s);
    return launch_v<true, false>(c, T, G, F, P, psi, one, none, s);
}

static int update_s_region(const ifos_geom *g, int region, const Box *box, const ifos_fields *f,
                           const ifos_model *m, const float *d_prof, const ifos_psi *psi_v, hipStream_t s,
                           const ifos_visco *visco = nullptr)
{
    DevGeom G = make_dev_geom(g);
    const SFields F = sfields(f, m);
    const Prof P = {d_prof, 2 * g->fw + 1};
    const Psi9 psi = psi9(psi_v);
    const Shots one = {1, 0, {0, 0, 0}};
    const MarchCfg &c = g_default_cfg;
    if (region == IFOS_REGION_CPML && !box) {
        const BoxList L = slab_list(g);
        const int64_t n = L.start[L.n];
        if (n == 0) return IFOS_OK;
        ProfScope ps_(IFOS_K_S_CPML, s);
        const dim3 grid((unsigned)((n + 255) / 256));
        if (visco && visco->L == 1)
            hipLaunchKernelGGL(k_s_slabs<1>, grid, dim3(256), 0, s, G, L, F, P, psi, viscof(g, visco));
        else if (visco)
            hipLaunchKernelGGL(k_s_slabs<2>, grid, dim3(256), 0, s, G, L, F, P, psi, viscof(g, visco));
        else
            hipLaunchKernelGGL(k_s_slabs<0>, grid, dim3(256), 0, s, G, L, F, P, psi, ViscoF{});
        return hip_check(hipGetLastError(), "k_s_slabs");
    }
    Box boxes[1];
    boxes[0] = box ? *box : region_box(g, region);
    const MTiles T = make_tiles(boxes, 1, c.ti, c.jc, 1, 0);
    const int cat = region == IFOS_REGION_ALL ? IFOS_K_S_FUSED : IFOS_K_S_INTERIOR;
    ProfScope ps_(cat, s);
    if (visco) {
        const ViscoF V = viscof(g, visco);
        if (visco->L == 1) {
            if (region == IFOS_REGION_INTERIOR) return launch_s<false, 1>(c, T, G, F, P, psi, one, s, V);
            return launch_s<true, 1>(c, T, G, F, P, psi, one, s, V);
        }
        if (region == IFOS_REGION_INTERIOR) return launch_s<false, 2>(c, T, G, F, P, psi, one, s, V);
        return launch_s<true, 2>(c, T, G, F, P, psi, one, s, V);
    }
    if (region == IFOS_REGION_INTERIOR) return launch_s<false>(c, T, G, F, P, psi, one, s);
    return launch_s<true>(c, T, G, F, P, psi, one, s);
}

}  // namespace ifos

using namespace ifos;

extern "C" int ifos_update_v(const ifos_geom *g, int region, const ifos_fields *f, const ifos_model *m,
                             const float *d_prof, const ifos_psi *psi_s, void *stream)
{
    if (!g || !f || !m) {
        set_error("ifos_update_v: null argument");
        return IFOS_EINVAL;
    }
    if (region != IFOS_REGION_INTERIOR && (!d_prof || !psi_s)) {
        set_error("ifos_update_v: CPML region needs profiles and psi arrays");
        return IFOS_EINVAL;
    }
    return update_v_region(g, region, nullptr, f, m, d_prof, psi_s, (hipStream_t)stream);
}

extern "C" int ifos_update_s(const ifos_geom *g, int region, const ifos_fields *f, const ifos_model *m,
                             const float *d_prof, const ifos_psi *psi_v, void *stream)
{
    if (!g || !f || !m) {
        set_error("ifos_update_s: null argument");
        return IFOS_EINVAL;
    }
    if (region != IFOS_REGION_INTERIOR && (!d_prof || !psi_v)) {
        set_error("ifos_update_s: CPML region needs profiles and psi arrays");
        return IFOS_EINVAL;
    }
    return update_s_region(g, region, nullptr, f, m, d_prof, psi_v, (hipStream_t)stream);
}

extern "C" int ifos_update_s_visco(const ifos_geom *g, int region, const ifos_fields *f, const ifos_model *m,
                                   const ifos_visco *visco, const float *d_prof, const ifos_psi *psi_v,
                                   void *stream)
{
    if (!g || !f || !m || !visco) {
        set_error("ifos_update_s_visco: null argument");
        return IFOS_EINVAL;
    }
    if (visco->L < 1 || visco->L > IFOS_MAX_L ||
        (double)visco->L * (double)ifos_field_elems(g) * 4.0 >= 4294967296.0) {
        set_error("ifos_update_s_visco: L=%d (1..%d relaxation mechanisms, L fields of memory variables per "
                  "component under 4 GiB)", visco->L, IFOS_MAX_L);
        return IFOS_EINVAL;
    }
    if (region != IFOS_REGION_INTERIOR && (!d_prof || !psi_v)) {
        set_error("ifos_update_s_visco: CPML region needs profiles and psi arrays");
        return IFOS_EINVAL;
    }
    return update_s_region(g, region, nullptr, f, m, d_prof, psi_v, (hipStream_t)stream, visco);
}

// box-restricted interior updates (the reference's update_v / update_s_elastic
// take the box as arguments, update_v_ssg.c:27-30)
extern "C" int ifos_update_v_box(const ifos_geom *g, const int box[6], const ifos_fields *f, const ifos_model *m,
                                 void *stream)
{
    if (!g || !f || !m || !box) return IFOS_EINVAL;
    Box R = {box[0], box[1], box[2], box[3], box[4], box[5]};
    if (R.i0 < 1 || R.i1 > g->nx || R.j0 < 1 || R.j1 > g->ny || R.k0 < 1 || R.k1 > g->nz) {
        set_error("ifos_update_v_box: box outside [1..N]^3");
        return IFOS_EINVAL;
    }
    if (R.i1 < R.i0 || R.j1 < R.j0 || R.k1 < R.k0) return IFOS_OK;
    return update_v_region(g, IFOS_REGION_INTERIOR, &R, f, m, nullptr, nullptr, (hipStream_t)stream);
}

extern "C" int ifos_update_s_box(const ifos_geom *g, const int box[6], const ifos_fields *f, const ifos_model *m,
                                 void *stream)
{
    if (!g || !f || !m || !box) return IFOS_EINVAL;
    Box R = {box[0], box[1], box[2], box[3], box[4], box[5]};
    if (R.i0 < 1 || R.i1 > g->nx || R.j0 < 1 || R.j1 > g->ny || R.k0 < 1 || R.k1 > g->nz) {
        set_error("ifos_update_s_box: box outside [1..N]^3");
        return IFOS_EINVAL;
    }
    if (R.i1 < R.i0 || R.j1 < R.j0 || R.k1 < R.k0) return IFOS_OK;
    return update_s_region(g, IFOS_REGION_INTERIOR, &R, f, m, nullptr, nullptr, (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// multi-shot launches of the batched path: split = 0 one CPML-aware launch
// over [1..N]^3; split = 1 the interior box without CPML code plus the six
// slabs with it (two launches)
namespace ifos {

int launch_v_shots(const ifos_geom *g, const ifos_fields *f, const ifos_model *m, const float *d_prof,
                   const ifos_psi *psi_s, const Shots &S, const Snap *snap, const MarchCfg &c, hipStream_t s)
{
    DevGeom G = make_dev_geom(g);
    const VFields F = vfields(f, m);
    const Prof P = {d_prof, 2 * g->fw + 1};
    const Psi9 psi = psi9(psi_s);
    const Snap N = snap ? *snap : Snap{nullptr, 0, 0, 0};
    ProfScope ps_(IFOS_K_V_FUSED, s);
    if (!c.split) {
        const Box all = region_box(g, IFOS_REGION_ALL);
        const MTiles T = make_tiles(&all, 1, c.ti, c.jc, S.n, c.group, c.kb, c.sb);
        return snap ? launch_v<true, true>(c, T, G, F, P, psi, S, N, s)
                    : launch_v<true, false>(c, T, G, F, P, psi, S, N, s);
    }
    const Box in = region_box(g, IFOS_REGION_INTERIOR);
    Box sl[6];
    const int nb = slab_boxes(g, sl);
    const MTiles Ti = make_tiles(&in, 1, c.ti, c.jc, S.n, c.group, c.kb, c.sb);
    const MTiles Ts = make_tiles(sl, nb, c.ti, c.jc, S.n, c.group, c.kb, c.sb);
    int rc = snap ? launch_v<false, true>(c, Ti, G, F, P, psi, S, N, s)
                  : launch_v<false, false>(c, Ti, G, F, P, psi, S, N, s);
    if (!rc)
        rc = snap ? launch_v<true, true>(c, Ts, G, F, P, psi, S, N, s)
                  : launch_v<true, false>(c, Ts, G, F, P, psi, S, N, s);
    return rc;
}

int launch_s_shots(const ifos_geom *g, const ifos_fields *f, const ifos_model *m, const float *d_prof,
                   const ifos_psi *psi_v, const Shots &S, const MarchCfg &c, hipStream_t s, const ViscoF *V)
{
    DevGeom G = make_dev_geom(g);
    const SFields F = sfields(f, m);
    const Prof P = {d_prof, 2 * g->fw + 1};
    const Psi9 psi = psi9(psi_v);
    ProfScope ps_(IFOS_K_S_FUSED, s);
    auto go = [&](const MTiles &T, bool cpml) {
        if (!V) return cpml ? launch_s<true>(c, T, G, F, P, psi, S, s) : launch_s<false>(c, T, G, F, P, psi, S, s);
        if (V->L == 1)
            return cpml ? launch_s<true, 1>(c, T, G, F, P, psi, S, s, *V)
                        : launch_s<false, 1>(c, T, G, F, P, psi, S, s, *V);
        return cpml ? launch_s<true, 2>(c, T, G, F, P, psi, S, s, *V) : launch_s<false, 2>(c, T, G, F, P, psi, S, s, *V);
    };
    if (!c.split) {
        const Box all = region_box(g, IFOS_REGION_ALL);
        return go(make_tiles(&all, 1, c.ti, c.jc, S.n, c.group, c.kb, c.sb), true);
    }
    const Box in = region_box(g, IFOS_REGION_INTERIOR);
    Box sl[6];
    const int nb = slab_boxes(g, sl);
    int rc = go(make_tiles(&in, 1, c.ti, c.jc, S.n, c.group, c.kb, c.sb), false);
    if (!rc) rc = go(make_tiles(sl, nb, c.ti, c.jc, S.n, c.group, c.kb, c.sb), true);
    return rc;
}

}  // namespace ifos

namespace ifos {
// gradient_F over [1..N]^3 for one frequency: f[12] point at that
// frequency's fields (forward Re xyz, Im xyz, adjoint Re xyz, Im xyz)
int launch_grad_march(const ifos_geom *g, const float *const f[12], float fd, const ifos_model *m,
                      float *const d_grad[3], hipStream_t s, bool hess)
{
    DevGeom G = make_dev_geom(g);
    GradArgs A;
    for (int n = 0; n < 12; ++n) A.f[n] = f[n];
    A.fd = fd;
    A.rho = m->rho;
    A.pi = m->pi;
    A.u = m->u;
    A.g1 = d_grad[0];
    A.g2 = d_grad[1];
    A.g3 = d_grad[2];
    const Box all = region_box(g, IFOS_REGION_ALL);
    const MTiles T = make_tiles(&all, 1, 4, 0, 1, 0);
    if (hess)
        hipLaunchKernelGGL((k_grad_march<4, GRAD_HESS>), dim3((unsigned)(T.per * 8)), dim3(64, 4), 0, s, G, T, A);
    else
        hipLaunchKernelGGL((k_grad_march<4, GRAD_F>), dim3((unsigned)(T.per * 8)), dim3(64, 4), 0, s, G, T, A);
    return hip_check(hipGetLastError(), "k_grad_march");
}

// one time sample of the time-domain gradient over [1..N]^3: uf, vf forward
// displacement and velocity, ub, vb adjoint displacement and velocity
int launch_td_sample(const ifos_geom *g, const float *const uf[3], const float *const vf[3],
                     const float *const ub[3], const float *const vb[3], double w, double *const acc[3],
                     hipStream_t s)
{
    DevGeom G = make_dev_geom(g);
    GradArgs A = {};
    for (int n = 0; n < 3; ++n) {
        A.f[n] = uf[n];
        A.f[3 + n] = vf[n];
        A.f[6 + n] = ub[n];
        A.f[9 + n] = vb[n];
        A.acc[n] = acc[n];
    }
    A.w = w;
    const Box all = region_box(g, IFOS_REGION_ALL);
    const MTiles T = make_tiles(&all, 1, 4, 0, 1, 0);
    { ProfScope ps_(IFOS_K_GRADIENT, s);
    hipLaunchKernelGGL((k_grad_march<4, GRAD_TD>), dim3((unsigned)(T.per * 8)), dim3(64, 4), 0, s, G, T, A);
    }
    return hip_check(hipGetLastError(), "k_grad_march td");
}
}  // namespace ifos

// ---------------------------------------------------------------------------
// Fused time step of the batched path: the stress update of step n (with the
// explosive sources of step n) and the velocity update of step n+1 in one
// march over j, so each wavefield array crosses HBM once per step.
//
// The fields are double-buffered: v^n and sigma^(n-1) are read from one set,
// sigma^n and v^(n+1) written to the other, so no block ever reads a value
// another block has already overwritten.  A workgroup of 64 x TI lanes covers
// k-columns [k0-2, k0+62) x i-rows [i0-2, i0+TI-2); every lane computes
// sigma^n at its own (i, k), the 60 x (TI-4) inner lanes own their cells
// (they store sigma^n and the CPML memory of the stress update, and compute
// v^(n+1)).  The five stress components the velocity update needs in the
// (i, k) plane pass through LDS (a ring of four planes); the j-neighbours stay
// in each lane's register queues.  Plane j of sigma^n is computed two planes
// ahead of plane j-2 of v^(n+1).  Halo planes, rows and columns outside
// [1..N]^3 hold zero stress (no free surface).
// CPML of the velocity derivatives reading the previous memory from `po`
// and, for the owner, writing the new memory to `pn` (else pn.p[0] == null)
__device__ inline void s_cpml2(const DevGeom &g, const Prof &P, const Psi9 &po, const Psi9 &pn, bool own, int j,
                               int i, int k, int hx, int hy, int hz, float &vxx, float &vyx, float &vzx,
                               float &vxy, float &vyy, float &vzy, float &vxz, float &vyz, float &vzz)
{
#define C2(n, a_, d, K, A_, B_)                       \
    do {                                              \
        const float p_ = (B_) * po.p[n][a_] + (A_) * (d); \
        if (own) pn.p[n][a_] = p_;                    \
        d = d / (K) + p_;                             \
    } while (0)
    if (hx) {
        const int64_t a = psi_x_at(g, j, hx, k);
        C2(0, a, vxx, P.K(0, hx), P.A(0, hx), P.B(0, hx));
        C2(1, a, vyx, P.KH(0, hx), P.AH(0, hx), P.BH(0, hx));
        C2(2, a, vzx, P.KH(0, hx), P.AH(0, hx), P.BH(0, hx));
    }
    if (hy) {
        const int64_t a = psi_y_at(g, hy, i, k);
        C2(3, a, vxy, P.KH(1, hy), P.AH(1, hy), P.BH(1, hy));
        C2(4, a, vyy, P.K(1, hy), P.A(1, hy), P.B(1, hy));
        C2(5, a, vzy, P.KH(1, hy), P.AH(1, hy), P.BH(1, hy));
    }
    if (hz) {
        const int64_t a = psi_z_at(g, j, i, hz);
        const int kax = (k <= g.fw) ? 1 : 2;
        C2(6, a, vxz, P.KH(kax, hz), P.AH(2, hz), P.BH(2, hz));
        C2(7, a, vyz, P.KH(kax, hz), P.AH(2, hz), P.BH(2, hz));
        C2(8, a, vzz, P.K(kax, hz), P.A(2, hz), P.B(2, hz));
    }
#undef C2
}

namespace ifos {
// absolute-plane buffer access: vcol is the byte offset of (-2, i-2, k-2)
#define LDJ(r, jj, di, dk) bload(r, vcol + ((dk) + 2) * 4, (uint32_t)((jj) + 2) * SJ + ((di) + 2) * SI)
#define STJ(v, r, jj) bstore(v, r, vcol + 8, (uint32_t)((jj) + 2) * SJ + 2 * SI)

template <int TI, bool SNAP>
__global__ __launch_bounds__(64 * TI) void k_vs_fused(DevGeom g, FTiles T, FusedArgs A, Prof P, Shots S, Snap N)
{
    __shared__ float lds[4][5][TI][64];  // sxx sxy sxz syz szz of sigma^n, plane slot j & 3
    const int b = blockIdx.x;
    const int L = (b & 7) * T.per + (b >> 3);
    if (L >= T.total) return;  // whole workgroup
    const int tiles = T.ntk * T.nti * T.ntj;
    const int shot = L / tiles;
    int t = L % tiles;
    const int tk = t % T.ntk;
    t /= T.ntk;
    const int ti = t % T.nti, tj = t / T.nti;
    const int lx = (int)threadIdx.x, ty = (int)threadIdx.y;
    const int k = tk * 60 - 1 + lx;        // k0 = 1 + 60 tk, lanes from k0 - 2
    const int i = ti * (TI - 4) - 1 + ty;  // i0 = 1 + (TI-4) ti, rows from i0 - 2
    const int jb = 1 + tj * T.jc, je = min(jb + T.jc - 1, g.ny);
    const bool col = k >= 1 && k <= g.nz && i >= 1 && i <= g.nx;  // (i, k) inside [1..N]^2
    const bool inner = col && lx >= 2 && lx < 62 && ty >= 2 && ty < TI - 2;
    const uint32_t SJ = (uint32_t)g.sj * 4u, SI = (uint32_t)g.si * 4u;
    const uint32_t vcol = col ? (uint32_t)((int64_t)i * g.si + k - 2 + IFOS_KOFF) * 4u : 0u;
    const float b1 = g.b1, b2 = g.b2;
    const float DT = g.dt;
    const int64_t so = (int64_t)shot * S.fstride;
    const rsrc_t rvx = make_rsrc(A.vx + so, g.fbytes), rvy = make_rsrc(A.vy + so, g.fbytes),
                 rvz = make_rsrc(A.vz + so, g.fbytes), rsxx = make_rsrc(A.sxx + so, g.fbytes),
                 rsyy = make_rsrc(A.syy + so, g.fbytes), rszz = make_rsrc(A.szz + so, g.fbytes),
                 rsxy = make_rsrc(A.sxy + so, g.fbytes), rsyz = make_rsrc(A.syz + so, g.fbytes),
                 rsxz = make_rsrc(A.sxz + so, g.fbytes);
    const rsrc_t rpi = make_rsrc(A.pi, g.fbytes), ru = make_rsrc(A.u, g.fbytes),
                 ruipjp = make_rsrc(A.uipjp, g.fbytes), rujpkp = make_rsrc(A.ujpkp, g.fbytes),
                 ruipkp = make_rsrc(A.uipkp, g.fbytes), rrip = make_rsrc(A.rip, g.fbytes),
                 rrjp = make_rsrc(A.rjp, g.fbytes), rrkp = make_rsrc(A.rkp, g.fbytes);
    const rsrc_t wvx = make_rsrc(A.nvx + so, g.fbytes), wvy = make_rsrc(A.nvy + so, g.fbytes),
                 wvz = make_rsrc(A.nvz + so, g.fbytes), wsxx = make_rsrc(A.nsxx + so, g.fbytes),
                 wsyy = make_rsrc(A.nsyy + so, g.fbytes), wszz = make_rsrc(A.nszz + so, g.fbytes),
                 wsxy = make_rsrc(A.nsxy + so, g.fbytes), wsyz = make_rsrc(A.nsyz + so, g.fbytes),
                 wsxz = make_rsrc(A.nsxz + so, g.fbytes);
    rsrc_t rring;
    if (SNAP) rring = make_rsrc(N.ring + (int64_t)shot * N.rstride, N.rbytes);
    Psi9 pvo = A.psv_old, pvn = A.psv_new, pss = A.pss;
    shift_psi(pvo, S, shot);
    shift_psi(pvn, S, shot);
    shift_psi(pss, S, shot);
    const int hx = col ? slab_h(i, g.nx, g.fw, 1) : 0;
    const int hz = col ? slab_h(k, g.nz, g.fw, 1) : 0;
    // explosive source of this shot, psource.c:40-50
    int si_ = -9, sj_ = -9, sk_ = -9;
    float amp = 0.0f;
    if (A.src) {
        const int *p = A.src + 4 * shot;
        if (p[3] == 1) {
            si_ = p[0];
            sj_ = p[1];
            sk_ = p[2];
            const float *sg = A.sig + (int64_t)shot * A.ldsig;
            const double den = 2.0 * g.dx * g.dy * g.dz;
            const int nt = A.nt;
            if (nt == 1) amp = (float)(sg[nt + 1] / den);
            if (nt > 1 && nt < A.ntall) amp = (float)((sg[nt + 1] - sg[nt - 1]) / den);
            if (nt == A.ntall) amp = (float)(-sg[nt - 1] / den);
        }
    }

    // v^n queues of the own column, planes j-2..j+2 at iteration j
    float qvx[5], qvy[5], qvz[5];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int p = jb - 4 + q;
        const bool ok = col && p >= -2;
        qvx[q] = ok ? LDJ(rvx, p, 0, 0) : 0.0f;
        qvy[q] = ok ? LDJ(rvy, p, 0, 0) : 0.0f;
        qvz[q] = ok ? LDJ(rvz, p, 0, 0) : 0.0f;
    }
    qvx[4] = qvy[4] = qvz[4] = 0.0f;
    // sigma^n queues of the own column: sxy, syz planes j-4..j, syy j-3..j
    float qxy[5] = {0, 0, 0, 0, 0}, qyz[5] = {0, 0, 0, 0, 0}, qyy[4] = {0, 0, 0, 0};

    for (int j = jb - 2; j <= je + 2; ++j) {
        // ---- phase A: sigma^n at plane j
        {
            const int p = j + 2;
            const bool ok = col && p <= g.ny + 2;
            qvx[4] = ok ? LDJ(rvx, p, 0, 0) : 0.0f;
            qvy[4] = ok ? LDJ(rvy, p, 0, 0) : 0.0f;
            qvz[4] = ok ? LDJ(rvz, p, 0, 0) : 0.0f;
        }
        float sn[6] = {0, 0, 0, 0, 0, 0};  // sxx syy szz sxy syz sxz
        if (col && j >= 1 && j <= g.ny) {
            const float vx_0 = qvx[2], vy_0 = qvy[2], vz_0 = qvz[2];
            const float vx_xa = LDJ(rvx, j, -2, 0), vx_xb = LDJ(rvx, j, -1, 0), vx_xd = LDJ(rvx, j, 1, 0);
            const float vx_zb = LDJ(rvx, j, 0, -1), vx_zc = LDJ(rvx, j, 0, 1), vx_zd = LDJ(rvx, j, 0, 2);
            const float vy_xb = LDJ(rvy, j, -1, 0), vy_xc = LDJ(rvy, j, 1, 0), vy_xd = LDJ(rvy, j, 2, 0);
            const float vy_zb = LDJ(rvy, j, 0, -1), vy_zc = LDJ(rvy, j, 0, 1), vy_zd = LDJ(rvy, j, 0, 2);
            const float vz_xb = LDJ(rvz, j, -1, 0), vz_xc = LDJ(rvz, j, 1, 0), vz_xd = LDJ(rvz, j, 2, 0);
            const float vz_za = LDJ(rvz, j, 0, -2), vz_zb = LDJ(rvz, j, 0, -1), vz_zd = LDJ(rvz, j, 0, 1);
            const float uipjp = LDJ(ruipjp, j, 0, 0), ujpkp = LDJ(rujpkp, j, 0, 0), uipkp = LDJ(ruipkp, j, 0, 0);
            const float pi = LDJ(rpi, j, 0, 0), u = LDJ(ru, j, 0, 0);
            const float sxx0 = LDJ(rsxx, j, 0, 0), syy0 = LDJ(rsyy, j, 0, 0), szz0 = LDJ(rszz, j, 0, 0);
            const float sxy0 = LDJ(rsxy, j, 0, 0), syz0 = LDJ(rsyz, j, 0, 0), sxz0 = LDJ(rsxz, j, 0, 0);
            float vxx = D4B(b1, b2, vx_xa, vx_xb, vx_0, vx_xd);
            float vxy = D4F(b1, b2, qvx[1], vx_0, qvx[3], qvx[4]);
            float vxz = D4F(b1, b2, vx_zb, vx_0, vx_zc, vx_zd);
            float vyx = D4F(b1, b2, vy_xb, vy_0, vy_xc, vy_xd);
            float vyy = D4B(b1, b2, qvy[0], qvy[1], vy_0, qvy[3]);
            float vyz = D4F(b1, b2, vy_zb, vy_0, vy_zc, vy_zd);
            float vzx = D4F(b1, b2, vz_xb, vz_0, vz_xc, vz_xd);
            float vzy = D4F(b1, b2, qvz[1], vz_0, qvz[3], qvz[4]);
            float vzz = D4B(b1, b2, vz_za, vz_zb, vz_0, vz_zd);
            div_spacing(g, vxx, vxy, vxz, vyx, vyy, vyz, vzx, vzy, vzz);
            const bool own = inner && j >= jb && j <= je;
            const int hy = slab_h(j, g.ny, g.fw, !g.free_surf);
            if (hx | hy | hz)
                s_cpml2(g, P, pvo, pvn, own, j, i, k, hx, hy, hz, vxx, vyx, vzx, vxy, vyy, vzy, vxz, vyz, vzz);
            // update_s_ssg_elastic.c:117-138
            const float fipjp = uipjp * DT, fjpkp = ujpkp * DT, fipkp = uipkp * DT;
            const float f = 2.0f * u;
            const float vxyyx = vxy + vyx, vyzzy = vyz + vzy, vxzzx = vxz + vzx;
            const float vxxyyzz = vxx + vyy + vzz;
            const float vyyzz = vyy + vzz, vxxzz = vxx + vzz, vxxyy = vxx + vyy;
            sn[3] = sxy0 + (fipjp * vxyyx);
            sn[4] = syz0 + (fjpkp * vyzzy);
            sn[5] = sxz0 + (fipkp * vxzzx);
            sn[0] = sxx0 + DT * ((pi * vxxyyzz) - (f * vyyzz));
            sn[1] = syy0 + DT * ((pi * vxxyyzz) - (f * vxxzz));
            sn[2] = szz0 + DT * ((pi * vxxyyzz) - (f * vxxyy));
            if (i == si_ && j == sj_ && k == sk_) {  // psource after the stress update
                sn[0] += amp;
                sn[1] += amp;
                sn[2] += amp;
            }
            if (own) {
                STJ(sn[3], wsxy, j);
                STJ(sn[4], wsyz, j);
                STJ(sn[5], wsxz, j);
                STJ(sn[0], wsxx, j);
                STJ(sn[1], wsyy, j);
                STJ(sn[2], wszz, j);
            }
        }
        const int slot = j & 3;
        lds[slot][0][ty][lx] = sn[0];
        lds[slot][1][ty][lx] = sn[3];
        lds[slot][2][ty][lx] = sn[5];
        lds[slot][3][ty][lx] = sn[4];
        lds[slot][4][ty][lx] = sn[2];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            qxy[q] = qxy[q + 1];
            qyz[q] = qyz[q + 1];
        }
        qxy[4] = sn[3];
        qyz[4] = sn[4];
#pragma unroll
        for (int q = 0; q < 3; ++q) qyy[q] = qyy[q + 1];
        qyy[3] = sn[1];
        __syncthreads();

        // ---- phase B: v^(n+1) at plane jv = j - 2
        const int jv = j - 2;
        if (inner && jv >= jb && jv <= je) {
            const int sl = jv & 3;
            const float(*X)[64] = lds[sl][0];
            const float(*XY)[64] = lds[sl][1];
            const float(*XZ)[64] = lds[sl][2];
            const float(*YZ)[64] = lds[sl][3];
            const float(*ZZ)[64] = lds[sl][4];
            const float sxy_0 = qxy[2], syz_0 = qyz[2], sxz_0 = XZ[ty][lx];
            float sxx_x = g.dtdx * D4F(b1, b2, X[ty - 1][lx], X[ty][lx], X[ty + 1][lx], X[ty + 2][lx]);
            float sxy_y = g.dtdy * D4B(b1, b2, qxy[0], qxy[1], sxy_0, qxy[3]);
            float sxz_z = g.dtdz * D4B(b1, b2, XZ[ty][lx - 2], XZ[ty][lx - 1], sxz_0, XZ[ty][lx + 1]);
            float syy_y = g.dtdy * D4F(b1, b2, qyy[0], qyy[1], qyy[2], qyy[3]);
            float sxy_x = g.dtdx * D4B(b1, b2, XY[ty - 2][lx], XY[ty - 1][lx], sxy_0, XY[ty + 1][lx]);
            float syz_z = g.dtdz * D4B(b1, b2, YZ[ty][lx - 2], YZ[ty][lx - 1], syz_0, YZ[ty][lx + 1]);
            float szz_z = g.dtdz * D4F(b1, b2, ZZ[ty][lx - 1], ZZ[ty][lx], ZZ[ty][lx + 1], ZZ[ty][lx + 2]);
            float sxz_x = g.dtdx * D4B(b1, b2, XZ[ty - 2][lx], XZ[ty - 1][lx], sxz_0, XZ[ty + 1][lx]);
            float syz_y = g.dtdy * D4B(b1, b2, qyz[0], qyz[1], syz_0, qyz[3]);
            const int hy = slab_h(jv, g.ny, g.fw, !g.free_surf);
            if (hx | hy | hz)
                v_cpml(g, P, pss, jv, i, k, hx, hy, hz, sxx_x, sxy_x, sxz_x, sxy_y, syy_y, syz_y, sxz_z, syz_z,
                       szz_z);
            const float rip = LDJ(rrip, jv, 0, 0), rjp = LDJ(rrjp, jv, 0, 0), rkp = LDJ(rrkp, jv, 0, 0);
            const float nvx = qvx[0] + (sxx_x + sxy_y + sxz_z) / rip;
            const float nvy = qvy[0] + (syy_y + sxy_x + syz_z) / rjp;
            const float nvz = qvz[0] + (szz_z + sxz_x + syz_y) / rkp;
            STJ(nvx, wvx, jv);
            STJ(nvy, wvy, jv);
            STJ(nvz, wvz, jv);
            if (SNAP) {
                const uint32_t o = (uint32_t)(jv + 2) * SJ + 2 * SI + N.soff;
                bstore(nvx, rring, vcol + 8, o);
                bstore(nvy, rring, vcol + 8, o + g.fbytes);
                bstore(nvz, rring, vcol + 8, o + 2 * g.fbytes);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            qvx[q] = qvx[q + 1];
            qvy[q] = qvy[q + 1];
            qvz[q] = qvz[q + 1];
        }
    }
}
#undef LDJ
#undef STJ

// one fused step for S.n shots (TI 8 or 16 i-rows per workgroup, jc j-planes
// per tile, <= 0 automatic)
int launch_vs_fused(const ifos_geom *g, const FusedArgs &A, const float *d_prof, const Shots &S, const Snap *snap,
                    int ti, int jc, hipStream_t s)
{
    DevGeom G = make_dev_geom(g);
    const Prof P = {d_prof, 2 * g->fw + 1};
    if (ti != 8 && ti != 16) {
        set_error("fused step: TI=%d not in {8,16}", ti);
        return IFOS_EINVAL;
    }
    FTiles T;
    T.ntk = (g->nz + 59) / 60;
    T.nti = (g->nx + ti - 5) / (ti - 4);
    if (jc <= 0) {
        jc = 64;
        while (jc > 8 && (int64_t)T.ntk * T.nti * ((g->ny + jc - 1) / jc) * S.n < 2048) jc /= 2;
    }
    T.jc = jc;
    T.ntj = (g->ny + jc - 1) / jc;
    T.ns = S.n;
    T.total = T.ntk * T.nti * T.ntj * S.n;
    T.per = (T.total + 7) / 8;
    const Snap N = snap ? *snap : Snap{nullptr, 0, 0, 0};
    const dim3 grid((unsigned)(T.per * 8));
    ProfScope ps_(IFOS_K_VS_FUSED, s);
    if (ti == 8) {
        if (snap) hipLaunchKernelGGL((k_vs_fused<8, true>), grid, dim3(64, 8), 0, s, G, T, A, P, S, N);
        else hipLaunchKernelGGL((k_vs_fused<8, false>), grid, dim3(64, 8), 0, s, G, T, A, P, S, N);
    } else {
        if (snap) hipLaunchKernelGGL((k_vs_fused<16, true>), grid, dim3(64, 16), 0, s, G, T, A, P, S, N);
        else hipLaunchKernelGGL((k_vs_fused<16, false>), grid, dim3(64, 16), 0, s, G, T, A, P, S, N);
    }
    return hip_check(hipGetLastError(), "k_vs_fused");
}
}  // namespace ifos

extern "C" int ifos_set_plain_tiles(const int cfg[IFOS_TILE_FIELDS])
{
    if (!cfg) return IFOS_EINVAL;
    if (ifos::check_tile_cfg(cfg, "ifos_set_plain_tiles")) return IFOS_EINVAL;
    ifos::g_default_cfg = ifos::MarchCfg{cfg[0], cfg[1], cfg[2], cfg[3], cfg[4], cfg[5], cfg[6], cfg[7]};
    return IFOS_OK;
}

extern "C" int ifos_get_plain_tiles(int cfg[IFOS_TILE_FIELDS])
{
    if (!cfg) return IFOS_EINVAL;
    const ifos::MarchCfg &c = ifos::g_default_cfg;
    const int v[IFOS_TILE_FIELDS] = {c.ti, c.jc, c.group, c.split, c.pipe, c.kb, c.sb, c.lds};
    for (int n = 0; n < IFOS_TILE_FIELDS; ++n) cfg[n] = v[n];
    return IFOS_OK;
}
