/*
 * ifos_oracle.h -- CPU restatement of the IFOS3D hot path (TEST INFRASTRUCTURE).
 *
 * This is the oracle the HIP library is checked against.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker.  The product path never calls it.
 *
 * Built twice from the same source (oracle/Makefile):
 *   liboracle_f32.so  ORACLE_REAL=float   -- "compat" precision: the exact
 *                     float/double mix of the reference C code, so results
 *                     are bit-identical to the compiled reference
 *                     (oracle/_ref) on the same inputs;
 *   liboracle_f64.so  ORACLE_REAL=double  -- every field and intermediate in
 *                     double: the FP64 oracle that FP32 errors are stated
 *                     against.
 *
 * Layout (the "canonical" layout, shared with the Python harness):
 *   every 3-D field is one contiguous block indexed [j][i][k] (k fastest,
 *   j = vertical, slowest) with a halo of 2 cells on every side:
 *   j in [-2, ny+2], i in [-2, nx+2], k in [-2, nz+2]; element (j,i,k) is at
 *   ((j+2)*(nx+5) + (i+2))*(nz+5) + (k+2).
 *   CPML memory arrays keep the reference's compact slab shapes
 *   (ifos3d.c:421-440), 0-based: psi_x[ny][2fw][nz], psi_y[2fw][nx][nz],
 *   psi_z[ny][nx][2fw].  CPML profile vectors are 1-based [1..2fw]
 *   (index 0 unused), as in the reference.
 */
#ifndef IFOS_ORACLE_H
#define IFOS_ORACLE_H

#include <stddef.h>

#ifndef ORACLE_REAL
#define ORACLE_REAL float
#endif
typedef ORACLE_REAL oreal;

#ifdef __cplusplus
extern "C" {
#endif

typedef struct og_grid {
    int nx, ny, nz;          /* local grid including CPML cells (NX, NY, NZ) */
    int fw;                  /* CPML width (FW) */
    int free_surf;           /* FREE_SURF */
    int fdcoeff;             /* 1 = Taylor, 2 = Holberg (FDCOEFF) */
    float dx, dy, dz, dt;    /* DX, DY, DZ, DT */
    int nt;                  /* NT */
    float time;              /* TIME = (NT-1)*DT, ifos3d.c:177 */
} og_grid;

/* number of elements of one canonical 3-D field */
int og_threads(void);
size_t og_field_size(const og_grid *g);
/* flat offset of (j,i,k) in a canonical field */
size_t og_idx(const og_grid *g, int j, int i, int k);

/* interior box, CPML_ini_elastic.c:26-71 (single rank: POS=0, NPROC=1).
 * box = {xb0, xb1, yb0, yb1, zb0, zb1} */
void og_interior_box(const og_grid *g, int box[6]);

/* CPML profiles, CPML_coeff.c:87-324.  Each output is a 1-based vector of
 * length 2*fw+1 (element 0 unused).  Order of the 18 kernel-visible outputs:
 * K_x a_x b_x K_x_half a_x_half b_x_half, then y, then z.  alpha_prime_* are
 * returned in alpha[6][2fw+1] (x, x_half, y, y_half, z, z_half). */
void og_cpml_coeff(const og_grid *g, float vppml, float fpml,
                   oreal *prof[18], oreal *alpha[6]);

/* single-rank matcopy (matcopy.c:31-307 with MPI_PROC_NULL neighbours):
 * halos of a model array become a periodic wrap of the local grid. */
void og_matcopy_wrap(const og_grid *g, oreal *m);

/* av_mat.c:28-80: harmonic mu averages and arithmetic rho (and tau)
 * averages on [1..N]^3.  taus/taus* may be NULL (L = 0). */
void og_av_mat(const og_grid *g, const oreal *rho, const oreal *u,
               const oreal *taus,
               oreal *uipjp, oreal *ujpkp, oreal *uipkp,
               oreal *tausipjp, oreal *tausjpkp, oreal *tausipkp,
               oreal *rjp, oreal *rkp, oreal *rip);

/* source description: one row of srcpos_loc (splitsrc.c:44-50) */
typedef struct og_source {
    int i, j, k;             /* grid indices (x, y vertical, z) */
    float tshift, fc, amp;   /* srcpos_loc[4..6] */
    int type;                /* srcpos_loc[7]: 1 explosive, 2-5 forces */
} og_source;

/* wavelet.c:47-90; signals is [nsrc][nt+2] indexed by nt (1-based). */
void og_wavelet(const og_grid *g, const og_source *src, int nsrc, int shape,
                float *signals);

/* update_v_ssg.c:27-556 (4th order): interior box, then body forces
 * (back = 0, :474-510) or raw adjoint traces (back = 1, :512-522).
 * For back = 1, srcs are the receiver positions and sig_x/y/z the
 * adjoint traces [ntr][nt+2] indexed by nt. */
void og_update_v(const og_grid *g, const int box[6], int nt,
                 oreal *vx, oreal *vy, oreal *vz,
                 const oreal *sxx, const oreal *syy, const oreal *szz,
                 const oreal *sxy, const oreal *syz, const oreal *sxz,
                 const oreal *rjp, const oreal *rkp, const oreal *rip,
                 const og_source *srcs, const float *sig_x,
                 const float *sig_y, const float *sig_z, int nsrc, int back);

/* update_v_ssg_CPML.c:30-403: the six slabs, psi recursion, and (back = 1)
 * the second adjoint injection. psi order: sxx_x sxy_x sxz_x sxy_y syy_y
 * syz_y sxz_z syz_z szz_z. */
void og_update_v_cpml(const og_grid *g, const int box[6], int nt,
                      oreal *vx, oreal *vy, oreal *vz,
                      const oreal *sxx, const oreal *syy, const oreal *szz,
                      const oreal *sxy, const oreal *syz, const oreal *sxz,
                      const oreal *rjp, const oreal *rkp, const oreal *rip,
                      const og_source *srcs, const float *sig_x,
                      const float *sig_y, const float *sig_z, int nsrc,
                      int back, oreal *const prof[18], oreal *psi[9]);

/* update_s_ssg_elastic.c:28-147 (4th order), interior box */
void og_update_s_elastic(const og_grid *g, const int box[6],
                         const oreal *vx, const oreal *vy, const oreal *vz,
                         oreal *sxx, oreal *syy, oreal *szz,
                         oreal *sxy, oreal *syz, oreal *sxz,
                         const oreal *pi, const oreal *u,
                         const oreal *uipjp, const oreal *ujpkp,
                         const oreal *uipkp);

/* update_s_ssg_CPML_elastic.c:30-505. psi order: vxx vyx vzx vxy vyy vzy
 * vxz vyz vzz. */
void og_update_s_cpml_elastic(const og_grid *g, const int box[6],
                              const oreal *vx, const oreal *vy,
                              const oreal *vz, oreal *sxx, oreal *syy,
                              oreal *szz, oreal *sxy, oreal *syz, oreal *sxz,
                              const oreal *pi, const oreal *u,
                              const oreal *uipjp, const oreal *ujpkp,
                              const oreal *uipkp, oreal *const prof[18],
                              oreal *psi[9]);

/* Viscoelastic stress update, 4th order (update_s_ssg.c:152-245 interior
 * box, update_s_ssg_CPML.c:30-772 slabs).  The reference hard-codes one
 * mechanism (update_s_ssg.c:110,211; read_par_json.c:268-282); L > 1
 * sums the L memory variables of each stress component where the
 * reference uses its one (before and after their update) and updates each
 * mechanism with its own eta, in mechanism order -- for L = 1 it is the
 * reference's arithmetic exactly.
 * r[6] = rxx ryy rzz rxy ryz rxz, each L canonical fields back to back
 * (mechanism l at r[n] + l * og_field_size); taus, taup cell values;
 * tausipjp/tausjpkp/tausipkp from og_av_mat; eta[l] = DT / relaxation time
 * of mechanism l (eta[1] of the reference for L = 1); L in 1..OG_MAX_L. */
#define OG_MAX_L 4
typedef struct og_visco {
    oreal *r[6];
    const oreal *taus, *taup, *tausipjp, *tausjpkp, *tausipkp;
    oreal eta[OG_MAX_L];
    int L;
} og_visco;

void og_update_s_visco(const og_grid *g, const int box[6], const oreal *vx,
                       const oreal *vy, const oreal *vz, oreal *sxx,
                       oreal *syy, oreal *szz, oreal *sxy, oreal *syz,
                       oreal *sxz, const oreal *pi, const oreal *u,
                       const oreal *uipjp, const oreal *ujpkp,
                       const oreal *uipkp, const og_visco *V);
void og_update_s_cpml_visco(const og_grid *g, const int box[6],
                            const oreal *vx, const oreal *vy,
                            const oreal *vz, oreal *sxx, oreal *syy,
                            oreal *szz, oreal *sxy, oreal *syz, oreal *sxz,
                            const oreal *pi, const oreal *u,
                            const oreal *uipjp, const oreal *ujpkp,
                            const oreal *uipkp, oreal *const prof[18],
                            oreal *psi[9], const og_visco *V);

/* psource.c:26-54 (explosive sources, type 1) */
void og_psource(const og_grid *g, int nt, oreal *sxx, oreal *syy, oreal *szz,
                const og_source *srcs, const float *signals, int nsrc);

/* surface_ssg_elastic.c:26-112 (4th order, free surface at j = 1) */
void og_surface_elastic(const og_grid *g, const oreal *u, const oreal *pi,
                        oreal *sxx, oreal *syy, oreal *szz, oreal *sxy,
                        oreal *syz, const oreal *vx, const oreal *vy,
                        const oreal *vz);

/* seismo_ssg.c:27-53 (SEISMO = 1): sections [ntr][ns+1], sample ins */
void og_seismo(const og_grid *g, int ins, int ntr, const int *rec_ijk,
               float *sec_x, float *sec_y, float *sec_z, int ns,
               const oreal *vx, const oreal *vy, const oreal *vz);

/* residual.c:27-56: time-reversed running integral of (syn - obs) into
 * diff [ntr][ns+1]; L2 accumulated in float as the reference does. */
void og_residual(const og_grid *g, const float *obs, const float *syn,
                 float *diff, int ntr, int ns, float *l2);

/* disc_fourier.c:28-62 on [1..N]^3. F arrays hold nf canonical fields each,
 * frequency l (1-based) at offset (l-1)*og_field_size. trig factors are
 * computed here from finv, exactly as the reference does. */
void og_discfourier(const og_grid *g, int nt, const oreal *vx,
                    const oreal *vy, const oreal *vz, oreal *fvx, oreal *fvy,
                    oreal *fvz, oreal *fivx, oreal *fivy, oreal *fivz,
                    const float *finv, int nf, int ntast, int back);

/* gradient_F.c:30-170: accumulates grad1..3 (vp, vs, rho) over [1..N]^3. */
void og_gradient_F(const og_grid *g, const oreal *fvx, const oreal *fvy,
                   const oreal *fvz, const oreal *fivx, const oreal *fivy,
                   const oreal *fivz, const oreal *bvx, const oreal *bvy,
                   const oreal *bvz, const oreal *bivx, const oreal *bivy,
                   const oreal *bivz, oreal *grad1, oreal *grad2,
                   oreal *grad3, const oreal *rho, const oreal *pi,
                   const oreal *u, const float *finv, int nf);

/* hess_F.c:28-192; receiver-side fields of receiver n, frequency l at
 * (n*nf + l - 1) * field size */
void og_hess_F(const og_grid *g, const oreal *fvx, const oreal *fvy,
               const oreal *fvz, const oreal *fivx, const oreal *fivy,
               const oreal *fivz, const oreal *bvx, const oreal *bvy,
               const oreal *bvz, const oreal *bivx, const oreal *bivy,
               const oreal *bivz, oreal *hess1, oreal *hess2, oreal *hess3,
               const oreal *rho, const oreal *pi, const oreal *u,
               const float *finv, int nf, int ntr_hess);

/* ntast, ifos3d.c:745-750 */
int og_ntast(float fmax, int tast, float dt);

/*
 * One gradient shot in the call order of ifos3d.c:712-975 (single rank,
 * elastic, CPML, NDT = 1, NDTSHIFT = 0, SEISMO = 1):
 *   forward NT steps (+ DFT every ntast steps, seismograms every step),
 *   residual against obs (if obs != NULL and want_adjoint),
 *   adjoint NT steps (+ DFT), gradient_F accumulated into grad1..3.
 * Inputs: model rho/pi/u canonical (halos are rebuilt by the single-rank
 * matcopy), one source, receivers rec_ijk[ntr][3] (i, j, k).
 * Outputs: syn_x/y/z [ntr][nt+1] (sample 1..nt), l2 (float sum), grads.
 * When want_adjoint == 0 only the forward is run.
 */
typedef struct og_shot_io {
    const oreal *rho, *pi, *u;
    og_source src;
    int source_shape;
    int ntr;
    const int *rec_ijk;
    const float *obs_x, *obs_y, *obs_z;     /* [ntr][nt+1] or NULL */
    const float *finv;
    int nf, tast;
    float vppml, fpml;
    int want_adjoint;
    float *syn_x, *syn_y, *syn_z;           /* [ntr][nt+1] */
    float *l2;
    oreal *grad1, *grad2, *grad3;           /* canonical, accumulated */
    /* optional: the frequency fields after the shot (6*nf canonical each),
     * NULL if not wanted */
    oreal *ffwd, *fadj;
    /* optional viscoelastic run (L = 1..OG_MAX_L mechanisms): taus, taup
     * canonical (cell values), eta[l] per mechanism; L = 0 (or NULL taus)
     * runs the elastic update of ifos3d.c */
    const oreal *taus, *taup;
    oreal eta[OG_MAX_L];
    int L;
    /* optional: the velocities vx, vy, vz after every step at the DFT point
     * of ifos3d.c (forward after psource, adjoint after the stress update),
     * [nt][3] canonical fields each, NULL if not wanted */
    oreal *vfwd, *vadj;
} og_shot_io;

int og_run_shot(const og_grid *g, og_shot_io *io);

/*
 * The time-domain gradient from the velocity histories of og_run_shot
 * (vfwd, vadj, [nt][3] canonical fields): the products of
 * gradient_F.c:122-132 summed over the time samples instead of the DFT
 * frequencies.  Forward sample n pairs with adjoint step m = nt - 1 - n
 * (the DFT times of disc_fourier.c:36-37), n = 1 .. nt - 2; lambda and mu
 * use the displacements integrated from the end of the record (forward:
 * uf(n) = sum over n' >= n of vf(n') DT; adjoint: ub(m) = sum over m' <= m
 * of vb(m') DT, both summed in oreal), rho the velocities; each sample is
 * weighted by DT and summed in double; the sums are added to grad1..3 with
 * the vp, vs, rho parametrisation of gradient_F.c:135-139 in double.
 */
void og_td_gradient(const og_grid *g, const oreal *vfwd, const oreal *vadj, const oreal *rho,
                    const oreal *pi, const oreal *u, oreal *grad1, oreal *grad2, oreal *grad3);

/*
 * The consistent (exact discrete) adjoint -- no counterpart in the reference,
 * whose adjoint is the forward scheme run with time-reversed, integrated
 * residuals (residual.c:43-53, injected twice: update_v_ssg.c:512-522 and
 * update_v_ssg_CPML.c:384-393) and whose gradient is the frequency-domain
 * estimate of gradient_F.c.  Here the misfit is
 *     J = 1/2 sum_{n=1..NT} sum_r sum_{x,y,z} (syn - obs)^2
 * of the very forward scheme of og_run_shot (elastic, 4th order, CPML, one
 * explosive source, no free surface), and the gradient is its exact
 * derivative: the transpose of every step (velocity update with its CPML
 * recursion, stress update with its CPML recursion, receiver sampling) run
 * backwards, the model derivatives of each update summed over the steps,
 * then carried through av_mat's averages and matcopy's halo copies to the
 * cells and to the (vp, vs, rho) parametrisation of gradient_F.c:135-139
 * (pi = rho vp^2, u = rho vs^2).  Returns 0, -1 on a NaN at the probe cell,
 * -2 for a configuration it does not cover.
 */
typedef struct og_adj_io {
    const oreal *rho, *pi, *u;              /* canonical; halos rebuilt by matcopy */
    og_source src;                          /* type 1 (explosive) */
    int source_shape;
    int ntr;
    const int *rec_ijk;
    const float *obs_x, *obs_y, *obs_z;     /* [ntr][nt+1] */
    float vppml, fpml;
    float *syn_x, *syn_y, *syn_z;           /* out [ntr][nt+1] */
    double *J;                              /* out */
    oreal *grad1, *grad2, *grad3;           /* dJ/dvp, dJ/dvs, dJ/drho, canonical, accumulated */
} og_adj_io;

int og_run_shot_adj(const og_grid *g, og_adj_io *io);

#ifdef __cplusplus
}
#endif
#endif
