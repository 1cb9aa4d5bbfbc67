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

// ---------------------------------------------------------------------------
// tile decode with XCD-aware renumbering: physical block b runs on XCD b % 8;
// logical tile L = (b % 8) * per + b / 8 gives each XCD a contiguous range.
struct Tiling {
    int ntk, nti, ntj, total, per;
    int jc;
    int ns;  // shots: total = ntk * nti * ntj * ns
};

__device__ inline bool decode_tile(const Tiling &T, int &tk, int &ti, int &tj, int &shot)
{
    int b = blockIdx.x;
    int L = (b & 7) * T.per + (b >> 3);
    if (L >= T.total) return false;
    shot = L % T.ns;
    L /= T.ns;
    tk = L % T.ntk;
    int r = L / T.ntk;
    ti = r % T.nti;
    tj = r / T.nti;
    return true;
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
#define ST(v, r) bstore(v, r, vo + 8, 2 * SJ + 2 * SI)

template <int TK, int TI, bool CPML, bool SNAP>
__global__ __launch_bounds__(TK *TI) void k_v_march(DevGeom g, Box R, Tiling T, VFields F, Prof P, Psi9 psi,
                                                   Shots S, Snap N)
{
    int tk, ti, tj, shot;
    if (!decode_tile(T, tk, ti, tj, shot)) return;
    const int k = R.k0 + tk * TK + (int)threadIdx.x;
    const int i = R.i0 + ti * TI + (int)threadIdx.y;
    if (k > R.k1 || i > R.i1) return;
    const int jb = R.j0 + tj * T.jc;
    const int je = min(jb + T.jc - 1, R.j1);
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

    // j queues: sxy and syz at j-2..j+1, syy at j-1..j+2
    float sxy_m2 = LD(rsxy, -2, 0, 0), sxy_m1 = LD(rsxy, -1, 0, 0), sxy_0 = LD(rsxy, 0, 0, 0),
          sxy_p1 = LD(rsxy, 1, 0, 0);
    float syz_m2 = LD(rsyz, -2, 0, 0), syz_m1 = LD(rsyz, -1, 0, 0), syz_0 = LD(rsyz, 0, 0, 0),
          syz_p1 = LD(rsyz, 1, 0, 0);
    float syy_m1 = LD(rsyy, -1, 0, 0), syy_0 = LD(rsyy, 0, 0, 0), syy_p1 = LD(rsyy, 1, 0, 0),
          syy_p2 = LD(rsyy, 2, 0, 0);

    for (int j = jb; j <= je; ++j, vo += SJ) {
        const float sxx_a = LD(rsxx, 0, -1, 0), sxx_b = LD(rsxx, 0, 0, 0), sxx_c = LD(rsxx, 0, 1, 0),
                    sxx_d = LD(rsxx, 0, 2, 0);
        const float sxy_a = LD(rsxy, 0, -2, 0), sxy_b = LD(rsxy, 0, -1, 0), sxy_d = LD(rsxy, 0, 1, 0);
        const float sxz_xa = LD(rsxz, 0, -2, 0), sxz_xb = LD(rsxz, 0, -1, 0), sxz_0 = LD(rsxz, 0, 0, 0),
                    sxz_xd = LD(rsxz, 0, 1, 0);
        const float sxz_za = LD(rsxz, 0, 0, -2), sxz_zb = LD(rsxz, 0, 0, -1), sxz_zd = LD(rsxz, 0, 0, 1);
        const float syz_za = LD(rsyz, 0, 0, -2), syz_zb = LD(rsyz, 0, 0, -1), syz_zd = LD(rsyz, 0, 0, 1);
        const float szz_a = LD(rszz, 0, 0, -1), szz_b = LD(rszz, 0, 0, 0), szz_c = LD(rszz, 0, 0, 1),
                    szz_d = LD(rszz, 0, 0, 2);
        const float rip = LD(rrip, 0, 0, 0), rjp = LD(rrjp, 0, 0, 0), rkp = LD(rrkp, 0, 0, 0);
        const float vx0 = LD(rvx, 0, 0, 0), vy0 = LD(rvy, 0, 0, 0), vz0 = LD(rvz, 0, 0, 0);
        float sxx_x = g.dtdx * D4F(b1, b2, sxx_a, sxx_b, sxx_c, sxx_d);
        float sxy_y = g.dtdy * D4B(b1, b2, sxy_m2, sxy_m1, sxy_0, sxy_p1);
        float sxz_z = g.dtdz * D4B(b1, b2, sxz_za, sxz_zb, sxz_0, sxz_zd);
        float syy_y = g.dtdy * D4F(b1, b2, syy_m1, syy_0, syy_p1, syy_p2);
        float sxy_x = g.dtdx * D4B(b1, b2, sxy_a, sxy_b, sxy_0, sxy_d);
        float syz_z = g.dtdz * D4B(b1, b2, syz_za, syz_zb, syz_0, syz_zd);
        float szz_z = g.dtdz * D4F(b1, b2, szz_a, szz_b, szz_c, szz_d);
        float sxz_x = g.dtdx * D4B(b1, b2, sxz_xa, sxz_xb, sxz_0, sxz_xd);
        float syz_y = g.dtdy * D4B(b1, b2, syz_m2, syz_m1, syz_0, syz_p1);
        if (CPML) {
            const int hy = slab_h(j, g.ny, g.fw, !g.free_surf);
            v_cpml(g, P, psi, j, i, k, hx, hy, hz, sxx_x, sxy_x, sxz_x, sxy_y, syy_y, syz_y, sxz_z, syz_z,
                   szz_z);
        }
        const float nvx = vx0 + (sxx_x + sxy_y + sxz_z) / rip;
        const float nvy = vy0 + (syy_y + sxy_x + syz_z) / rjp;
        const float nvz = vz0 + (szz_z + sxz_x + syz_y) / rkp;
        if (j < je) {  // next plane's queue operands, issued before this plane's stores
            sxy_m2 = sxy_m1; sxy_m1 = sxy_0; sxy_0 = sxy_p1; sxy_p1 = LD(rsxy, 2, 0, 0);
            syz_m2 = syz_m1; syz_m1 = syz_0; syz_0 = syz_p1; syz_p1 = LD(rsyz, 2, 0, 0);
            syy_m1 = syy_0; syy_0 = syy_p1; syy_p1 = syy_p2; syy_p2 = LD(rsyy, 3, 0, 0);
        }
        ST(nvx, rvx);
        ST(nvy, rvy);
        ST(nvz, rvz);
        if (SNAP) {
            bstore(nvx, rring, vo + 8, 2 * SJ + 2 * SI + N.soff);
            bstore(nvy, rring, vo + 8, 2 * SJ + 2 * SI + N.soff + g.fbytes);
            bstore(nvz, rring, vo + 8, 2 * SJ + 2 * SI + N.soff + 2 * g.fbytes);
        }
    }
}

template <int TK, int TI, bool CPML>
__global__ __launch_bounds__(TK *TI) void k_s_march(DevGeom g, Box R, Tiling T, SFields F, Prof P, Psi9 psi,
                                                   Shots S)
{
    int tk, ti, tj, shot;
    if (!decode_tile(T, tk, ti, tj, shot)) return;
    const int k = R.k0 + tk * TK + (int)threadIdx.x;
    const int i = R.i0 + ti * TI + (int)threadIdx.y;
    if (k > R.k1 || i > R.i1) return;
    const int jb = R.j0 + tj * T.jc;
    const int je = min(jb + T.jc - 1, R.j1);
    const uint32_t SJ = (uint32_t)g.sj * 4u, SI = (uint32_t)g.si * 4u;
    const float b1 = g.b1, b2 = g.b2;
    const float DX = g.dx, DY = g.dy, DZ = g.dz, DT = g.dt;
    const int64_t so = (int64_t)shot * S.fstride;
    const rsrc_t rvx = make_rsrc(F.vx + so, g.fbytes), rvy = make_rsrc(F.vy + so, g.fbytes),
                 rvz = make_rsrc(F.vz + so, g.fbytes), rsxx = make_rsrc(F.sxx + so, g.fbytes),
                 rsyy = make_rsrc(F.syy + so, g.fbytes), rszz = make_rsrc(F.szz + so, g.fbytes),
                 rsxy = make_rsrc(F.sxy + so, g.fbytes), rsyz = make_rsrc(F.syz + so, g.fbytes),
                 rsxz = make_rsrc(F.sxz + so, g.fbytes), rpi = make_rsrc(F.pi, g.fbytes), ru = make_rsrc(F.u, g.fbytes),
                 ruipjp = make_rsrc(F.uipjp, g.fbytes), rujpkp = make_rsrc(F.ujpkp, g.fbytes),
                 ruipkp = make_rsrc(F.uipkp, g.fbytes);
    if (CPML) shift_psi(psi, S, shot);
    uint32_t vo = (uint32_t)((int64_t)jb * g.sj + (int64_t)i * g.si + k + IFOS_KOFF - 2) * 4u;
    int hx = 0;
    if (CPML) hx = slab_h(i, g.nx, g.fw, 1);
    const int hz = CPML ? slab_h(k, g.nz, g.fw, 1) : 0;

    // j queues: vx, vz at j-1..j+2, vy at j-2..j+1
    float vx_m1 = LD(rvx, -1, 0, 0), vx_0 = LD(rvx, 0, 0, 0), vx_p1 = LD(rvx, 1, 0, 0), vx_p2 = LD(rvx, 2, 0, 0);
    float vz_m1 = LD(rvz, -1, 0, 0), vz_0 = LD(rvz, 0, 0, 0), vz_p1 = LD(rvz, 1, 0, 0), vz_p2 = LD(rvz, 2, 0, 0);
    float vy_m2 = LD(rvy, -2, 0, 0), vy_m1 = LD(rvy, -1, 0, 0), vy_0 = LD(rvy, 0, 0, 0), vy_p1 = LD(rvy, 1, 0, 0);

    for (int j = jb; j <= je; ++j, vo += SJ) {
        const float vx_xa = LD(rvx, 0, -2, 0), vx_xb = LD(rvx, 0, -1, 0), vx_xd = LD(rvx, 0, 1, 0);
        const float vx_zb = LD(rvx, 0, 0, -1), vx_zc = LD(rvx, 0, 0, 1), vx_zd = LD(rvx, 0, 0, 2);
        const float vy_xb = LD(rvy, 0, -1, 0), vy_xc = LD(rvy, 0, 1, 0), vy_xd = LD(rvy, 0, 2, 0);
        const float vy_zb = LD(rvy, 0, 0, -1), vy_zc = LD(rvy, 0, 0, 1), vy_zd = LD(rvy, 0, 0, 2);
        const float vz_xb = LD(rvz, 0, -1, 0), vz_xc = LD(rvz, 0, 1, 0), vz_xd = LD(rvz, 0, 2, 0);
        const float vz_za = LD(rvz, 0, 0, -2), vz_zb = LD(rvz, 0, 0, -1), vz_zd = LD(rvz, 0, 0, 1);
        const float uipjp = LD(ruipjp, 0, 0, 0), ujpkp = LD(rujpkp, 0, 0, 0), uipkp = LD(ruipkp, 0, 0, 0);
        const float pi = LD(rpi, 0, 0, 0), u = LD(ru, 0, 0, 0);
        const float sxx0 = LD(rsxx, 0, 0, 0), syy0 = LD(rsyy, 0, 0, 0), szz0 = LD(rszz, 0, 0, 0);
        const float sxy0 = LD(rsxy, 0, 0, 0), syz0 = LD(rsyz, 0, 0, 0), sxz0 = LD(rsxz, 0, 0, 0);
        float vxx = D4B(b1, b2, vx_xa, vx_xb, vx_0, vx_xd) / DX;
        float vxy = D4F(b1, b2, vx_m1, vx_0, vx_p1, vx_p2) / DY;
        float vxz = D4F(b1, b2, vx_zb, vx_0, vx_zc, vx_zd) / DZ;
        float vyx = D4F(b1, b2, vy_xb, vy_0, vy_xc, vy_xd) / DX;
        float vyy = D4B(b1, b2, vy_m2, vy_m1, vy_0, vy_p1) / DY;
        float vyz = D4F(b1, b2, vy_zb, vy_0, vy_zc, vy_zd) / DZ;
        float vzx = D4F(b1, b2, vz_xb, vz_0, vz_xc, vz_xd) / DX;
        float vzy = D4F(b1, b2, vz_m1, vz_0, vz_p1, vz_p2) / DY;
        float vzz = D4B(b1, b2, vz_za, vz_zb, vz_0, vz_zd) / DZ;
        if (CPML) {
            const int hy = slab_h(j, g.ny, g.fw, !g.free_surf);
            s_cpml(g, P, psi, j, i, k, hx, hy, hz, vxx, vyx, vzx, vxy, vyy, vzy, vxz, vyz, vzz);
        }
        // update_s_ssg_elastic.c:117-138
        const float fipjp = uipjp * DT, fjpkp = ujpkp * DT, fipkp = uipkp * DT;
        const float f = 2.0f * u;
        const float vxyyx = vxy + vyx, vyzzy = vyz + vzy, vxzzx = vxz + vzx;
        const float vxxyyzz = vxx + vyy + vzz;
        const float vyyzz = vyy + vzz, vxxzz = vxx + vzz, vxxyy = vxx + vyy;
        const float nsxy = sxy0 + (fipjp * vxyyx);
        const float nsyz = syz0 + (fjpkp * vyzzy);
        const float nsxz = sxz0 + (fipkp * vxzzx);
        const float nsxx = sxx0 + DT * ((pi * vxxyyzz) - (f * vyyzz));
        const float nsyy = syy0 + DT * ((pi * vxxyyzz) - (f * vxxzz));
        const float nszz = szz0 + DT * ((pi * vxxyyzz) - (f * vxxyy));
        if (j < je) {
            vx_m1 = vx_0; vx_0 = vx_p1; vx_p1 = vx_p2; vx_p2 = LD(rvx, 3, 0, 0);
            vz_m1 = vz_0; vz_0 = vz_p1; vz_p1 = vz_p2; vz_p2 = LD(rvz, 3, 0, 0);
            vy_m2 = vy_m1; vy_m1 = vy_0; vy_0 = vy_p1; vy_p1 = LD(rvy, 2, 0, 0);
        }
        ST(nsxy, rsxy);
        ST(nsyz, rsyz);
        ST(nsxz, rsxz);
        ST(nsxx, rsxx);
        ST(nsyy, rsyy);
        ST(nszz, rszz);
    }
}
#undef LD
#undef ST

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

__global__ __launch_bounds__(256) void k_s_slabs(DevGeom g, BoxList L, SFields F, Prof P, Psi9 psi)
{
    int j, i, k;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (!slab_cell(L, t, j, i, k)) return;
    const int64_t c = at(g, j, i, k), sj = g.sj, si = g.si;
    const float b1 = g.b1, b2 = g.b2;
    const float DX = g.dx, DY = g.dy, DZ = g.dz;
    float vxx = D4B(b1, b2, F.vx[c - 2 * si], F.vx[c - si], F.vx[c], F.vx[c + si]) / DX;
    float vxy = D4F(b1, b2, F.vx[c - sj], F.vx[c], F.vx[c + sj], F.vx[c + 2 * sj]) / DY;
    float vxz = D4F(b1, b2, F.vx[c - 1], F.vx[c], F.vx[c + 1], F.vx[c + 2]) / DZ;
    float vyx = D4F(b1, b2, F.vy[c - si], F.vy[c], F.vy[c + si], F.vy[c + 2 * si]) / DX;
    float vyy = D4B(b1, b2, F.vy[c - 2 * sj], F.vy[c - sj], F.vy[c], F.vy[c + sj]) / DY;
    float vyz = D4F(b1, b2, F.vy[c - 1], F.vy[c], F.vy[c + 1], F.vy[c + 2]) / DZ;
    float vzx = D4F(b1, b2, F.vz[c - si], F.vz[c], F.vz[c + si], F.vz[c + 2 * si]) / DX;
    float vzy = D4F(b1, b2, F.vz[c - sj], F.vz[c], F.vz[c + sj], F.vz[c + 2 * sj]) / DY;
    float vzz = D4B(b1, b2, F.vz[c - 2], F.vz[c - 1], F.vz[c], F.vz[c + 1]) / DZ;
    const int hx = slab_h(i, g.nx, g.fw, 1);
    const int hy = slab_h(j, g.ny, g.fw, !g.free_surf);
    const int hz = slab_h(k, g.nz, g.fw, 1);
    s_cpml(g, P, psi, j, i, k, hx, hy, hz, vxx, vyx, vzx, vxy, vyy, vzy, vxz, vyz, vzz);
    s_apply(g, F, c, vxx, vxy, vxz, vyx, vyy, vyz, vzx, vzy, vzz);
}

// ---------------------------------------------------------------------------
// host launchers
static const int TK = 64, TI = 4;

static Tiling make_tiling(const Box &R, int tk, int ti, int jc, int ns = 1)
{
    Tiling T;
    T.ntk = (R.k1 - R.k0 + 1 + tk - 1) / tk;
    T.nti = (R.i1 - R.i0 + 1 + ti - 1) / ti;
    T.ntj = (R.j1 - R.j0 + 1 + jc - 1) / jc;
    T.ns = ns;
    T.total = T.ntk * T.nti * T.ntj * ns;
    T.per = (T.total + 7) / 8;
    T.jc = jc;
    return T;
}

static int pick_jc(const Box &R, int ns = 1)
{
    // enough blocks to give every SIMD several waves: aim for >= 4096 tiles
    const int nk = (R.k1 - R.k0 + 1 + TK - 1) / TK, ni = (R.i1 - R.i0 + 1 + TI - 1) / TI;
    const int ny = R.j1 - R.j0 + 1;
    int jc = 64;
    while (jc > 8 && (int64_t)nk * ni * ((ny + jc - 1) / jc) * ns < 4096) jc /= 2;
    return jc;
}

static BoxList slab_boxes(const ifos_geom *g)
{
    BoxList L;
    const int fw = g->fw, nx = g->nx, ny = g->ny, nz = g->nz;
    const int xb0 = fw + 1, xb1 = nx - fw, yb0 = g->free_surf ? 1 : fw + 1, yb1 = ny - fw;
    Box all[6] = {
        {1, fw, 1, ny, 1, nz},                 // x low   (update_v_ssg_CPML.c:70)
        {xb1 + 1, nx, 1, ny, 1, nz},           // x high  (:142)
        {xb0, xb1, 1, fw, 1, nz},              // y low   (:215, no free surface)
        {xb0, xb1, yb1 + 1, ny, 1, nz},        // y high  (:267)
        {xb0, xb1, yb0, yb1, 1, fw},           // z low   (:320)
        {xb0, xb1, yb0, yb1, nz - fw + 1, nz}, // z high  (:352)
    };
    L.n = 0;
    L.start[0] = 0;
    for (int b = 0; b < 6; ++b) {
        if (b == 2 && g->free_surf) continue;
        const Box &B = all[b];
        if (B.i1 < B.i0 || B.j1 < B.j0 || B.k1 < B.k0) continue;
        L.b[L.n] = B;
        L.start[L.n + 1] = L.start[L.n] + (int64_t)(B.i1 - B.i0 + 1) * (B.j1 - B.j0 + 1) * (B.k1 - B.k0 + 1);
        ++L.n;
    }
    for (int b = L.n + 1; b < 7; ++b) L.start[b] = L.start[L.n];
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
    DevGeom G = make_dev_geom(g);
    VFields F = {f->vx, f->vy, f->vz, f->sxx, f->syy, f->szz, f->sxy, f->syz, f->sxz, m->rip, m->rjp, m->rkp};
    Prof P = {d_prof, 2 * g->fw + 1};
    Psi9 psi = {};
    if (psi_s)
        for (int n = 0; n < 9; ++n) psi.p[n] = psi_s->p[n];
    hipStream_t s = (hipStream_t)stream;
    if (region == IFOS_REGION_CPML) {
        BoxList L = slab_boxes(g);
        int64_t n = L.start[L.n];
        if (n == 0) return IFOS_OK;
        ProfScope ps_(IFOS_K_V_CPML, s);
        hipLaunchKernelGGL(k_v_slabs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, G, L, F, P, psi);
        return hip_check(hipGetLastError(), "k_v_slabs");
    }
    Box R = region_box(g, region);
    if (R.i1 < R.i0 || R.j1 < R.j0 || R.k1 < R.k0) return IFOS_OK;
    Tiling T = make_tiling(R, TK, TI, pick_jc(R));
    dim3 grid((unsigned)(T.per * 8)), block(TK, TI);
    const Shots one = {1, 0, {0, 0, 0}};
    const Snap none = {nullptr, 0, 0, 0};
    ProfScope ps_(region == IFOS_REGION_ALL ? IFOS_K_V_FUSED : IFOS_K_V_INTERIOR, s);
    if (region == IFOS_REGION_ALL)
        hipLaunchKernelGGL((k_v_march<TK, TI, true, false>), grid, block, 0, s, G, R, T, F, P, psi, one, none);
    else
        hipLaunchKernelGGL((k_v_march<TK, TI, false, false>), grid, block, 0, s, G, R, T, F, P, psi, one, none);
    return hip_check(hipGetLastError(), "k_v_march");
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
    DevGeom G = make_dev_geom(g);
    SFields F = {f->vx, f->vy, f->vz, f->sxx, f->syy, f->szz, f->sxy, f->syz, f->sxz,
                 m->pi, m->u, m->uipjp, m->ujpkp, m->uipkp};
    Prof P = {d_prof, 2 * g->fw + 1};
    Psi9 psi = {};
    if (psi_v)
        for (int n = 0; n < 9; ++n) psi.p[n] = psi_v->p[n];
    hipStream_t s = (hipStream_t)stream;
    if (region == IFOS_REGION_CPML) {
        BoxList L = slab_boxes(g);
        int64_t n = L.start[L.n];
        if (n == 0) return IFOS_OK;
        ProfScope ps_(IFOS_K_S_CPML, s);
        hipLaunchKernelGGL(k_s_slabs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, G, L, F, P, psi);
        return hip_check(hipGetLastError(), "k_s_slabs");
    }
    Box R = region_box(g, region);
    if (R.i1 < R.i0 || R.j1 < R.j0 || R.k1 < R.k0) return IFOS_OK;
    Tiling T = make_tiling(R, TK, TI, pick_jc(R));
    dim3 grid((unsigned)(T.per * 8)), block(TK, TI);
    const Shots one = {1, 0, {0, 0, 0}};
    ProfScope ps_(region == IFOS_REGION_ALL ? IFOS_K_S_FUSED : IFOS_K_S_INTERIOR, s);
    if (region == IFOS_REGION_ALL)
        hipLaunchKernelGGL((k_s_march<TK, TI, true>), grid, block, 0, s, G, R, T, F, P, psi, one);
    else
        hipLaunchKernelGGL((k_s_march<TK, TI, false>), grid, block, 0, s, G, R, T, F, P, psi, one);
    return hip_check(hipGetLastError(), "k_s_march");
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
    DevGeom G = make_dev_geom(g);
    VFields F = {f->vx, f->vy, f->vz, f->sxx, f->syy, f->szz, f->sxy, f->syz, f->sxz, m->rip, m->rjp, m->rkp};
    Prof P = {nullptr, 2 * g->fw + 1};
    Psi9 psi = {};
    hipStream_t s = (hipStream_t)stream;
    Tiling T = make_tiling(R, TK, TI, pick_jc(R));
    const Shots one = {1, 0, {0, 0, 0}};
    const Snap none = {nullptr, 0, 0, 0};
    ProfScope ps_(IFOS_K_V_INTERIOR, s);
    hipLaunchKernelGGL((k_v_march<TK, TI, false, false>), dim3((unsigned)(T.per * 8)), dim3(TK, TI), 0, s, G, R, T,
                       F, P, psi, one, none);
    return hip_check(hipGetLastError(), "k_v_march");
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
    DevGeom G = make_dev_geom(g);
    SFields F = {f->vx, f->vy, f->vz, f->sxx, f->syy, f->szz, f->sxy, f->syz, f->sxz,
                 m->pi, m->u, m->uipjp, m->ujpkp, m->uipkp};
    Prof P = {nullptr, 2 * g->fw + 1};
    Psi9 psi = {};
    hipStream_t s = (hipStream_t)stream;
    Tiling T = make_tiling(R, TK, TI, pick_jc(R));
    const Shots one = {1, 0, {0, 0, 0}};
    ProfScope ps_(IFOS_K_S_INTERIOR, s);
    hipLaunchKernelGGL((k_s_march<TK, TI, false>), dim3((unsigned)(T.per * 8)), dim3(TK, TI), 0, s, G, R, T, F, P,
                       psi, one);
    return hip_check(hipGetLastError(), "k_s_march");
}

// ---------------------------------------------------------------------------
// multi-shot fused launches of the batched path
namespace ifos {

int launch_v_shots(const ifos_geom *g, const ifos_fields *f, const ifos_model *m, const float *d_prof,
                   const ifos_psi *psi_s, const Shots &S, const Snap *snap, hipStream_t s)
{
    DevGeom G = make_dev_geom(g);
    VFields F = {f->vx, f->vy, f->vz, f->sxx, f->syy, f->szz, f->sxy, f->syz, f->sxz, m->rip, m->rjp, m->rkp};
    Prof P = {d_prof, 2 * g->fw + 1};
    Psi9 psi;
    for (int n = 0; n < 9; ++n) psi.p[n] = psi_s->p[n];
    Box R = region_box(g, IFOS_REGION_ALL);
    Tiling T = make_tiling(R, TK, TI, pick_jc(R, S.n), S.n);
    dim3 grid((unsigned)(T.per * 8)), block(TK, TI);
    ProfScope ps_(IFOS_K_V_FUSED, s);
    if (snap)
        hipLaunchKernelGGL((k_v_march<TK, TI, true, true>), grid, block, 0, s, G, R, T, F, P, psi, S, *snap);
    else
        hipLaunchKernelGGL((k_v_march<TK, TI, true, false>), grid, block, 0, s, G, R, T, F, P, psi, S,
                           Snap{nullptr, 0, 0, 0});
    return hip_check(hipGetLastError(), "k_v_march (shots)");
}

int launch_s_shots(const ifos_geom *g, const ifos_fields *f, const ifos_model *m, const float *d_prof,
                   const ifos_psi *psi_v, const Shots &S, hipStream_t s)
{
    DevGeom G = make_dev_geom(g);
    SFields F = {f->vx, f->vy, f->vz, f->sxx, f->syy, f->szz, f->sxy, f->syz, f->sxz,
                 m->pi, m->u, m->uipjp, m->ujpkp, m->uipkp};
    Prof P = {d_prof, 2 * g->fw + 1};
    Psi9 psi;
    for (int n = 0; n < 9; ++n) psi.p[n] = psi_v->p[n];
    Box R = region_box(g, IFOS_REGION_ALL);
    Tiling T = make_tiling(R, TK, TI, pick_jc(R, S.n), S.n);
    dim3 grid((unsigned)(T.per * 8)), block(TK, TI);
    ProfScope ps_(IFOS_K_S_FUSED, s);
    hipLaunchKernelGGL((k_s_march<TK, TI, true>), grid, block, 0, s, G, R, T, F, P, psi, S);
    return hip_check(hipGetLastError(), "k_s_march (shots)");
}

}  // namespace ifos
