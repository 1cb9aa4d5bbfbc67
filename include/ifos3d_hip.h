/*
 * ifos3d_hip.h -- native C ABI of the MI355X (gfx950) IFOS3D hot-path library
 * (libifos3d_hip.so).
 *
 * Device pointers, plain sizes, one HIP stream per call (passed as void*).
 * The reference-compatible entry points with the IFOS3D prototypes
 * (update_v, update_s_elastic, ..., gradient_F) live in ifos3d_dropin.h and
 * are built on top of these.
 *
 * Device layout of a 3-D field ("device layout"):
 *   j (vertical, slowest) in [-2, ny+2], i in [-2, nx+2], k (fastest) in
 *   [-2, nz+2]; element (j,i,k) at ((j+2)*(nx+5) + (i+2))*pitch + k + IFOS_KOFF.
 *   pitch is a multiple of 32 floats (128 B) and IFOS_KOFF puts k = 1 on a
 *   128-B boundary, so every interior row starts a cache line.
 *   This is the reference's NR layout (util.c:250-288: one block, k fastest,
 *   then i, then j) with a uniform 2-cell halo and a padded k pitch.
 * CPML memory arrays keep the reference's compact slab shapes
 *   (ifos3d.c:421-440): psi_x [ny][2fw][nz], psi_y [2fw][nx][nz],
 *   psi_z [ny][nx][2fw], dense, 0-based.
 * CPML profiles: one device array of 18*(2fw+1) floats, per axis
 *   (x, y, z) the six 1-based vectors K a b K_half a_half b_half
 *   (CPML_coeff.c:87-324), element 0 of each unused.
 *
 * Errors: every function returns 0 on success and a negative IFOS_E* code
 * otherwise; ifos_last_error() gives the message.  Nothing aborts.
 */
#ifndef IFOS3D_HIP_H
#define IFOS3D_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IFOS_KOFF 31

enum {
    IFOS_OK = 0,
    IFOS_EINVAL = -1,   /* bad argument / unsupported configuration */
    IFOS_EHIP = -2,     /* a HIP runtime call failed */
    IFOS_ENAN = -3,     /* simulation unstable: NaN at the probe cell */
    IFOS_ENOMEM = -4,
    IFOS_EUNSTABLE = -5, /* DT above the stability limit (checkfd_ssg.c:505-508) */
};

/* Regions of the velocity/stress updates.  INTERIOR is the box of
 * CPML_ini_elastic.c:26-71, CPML its complement in [1..N]^3 (the six
 * slabs of update_v_ssg_CPML.c / update_s_ssg_CPML_elastic.c), ALL both in
 * one fused pass. */
enum { IFOS_REGION_INTERIOR = 1, IFOS_REGION_CPML = 2, IFOS_REGION_ALL = 3 };

/* DFT accumulation precision */
enum {
    IFOS_DFT_COMPAT = 0, /* F = float(F + double(v)*trig): disc_fourier.c:48-54 */
    IFOS_DFT_KAHAN = 1,  /* compensated FP32 sum of the double product */
};

typedef struct ifos_geom {
    int nx, ny, nz;      /* local grid including CPML cells (NX, NY, NZ) */
    int fw;              /* CPML width (FW) */
    int free_surf;       /* FREE_SURF */
    int fdcoeff;         /* 1 Taylor, 2 Holberg (FDCOEFF); FDORDER is 4 */
    float dx, dy, dz, dt;
    int nt;              /* NT */
    float time;          /* TIME = (NT-1)*DT, ifos3d.c:177 */
    int pitch;           /* floats per (j,i) row, set by ifos_geom_init */
} ifos_geom;

/* Fills g (pitch, time) and validates the configuration. */
int ifos_geom_init(ifos_geom *g, int nx, int ny, int nz, int fw, float dx,
                   float dy, float dz, float dt, int nt, int fdcoeff,
                   int free_surf);
/* elements (floats) of one device-layout field */
size_t ifos_field_elems(const ifos_geom *g);
/* element offset of (j,i,k) */
size_t ifos_offset(const ifos_geom *g, int j, int i, int k);
/* elements of psi array n (0-2 x slab, 3-5 y slab, 6-8 z slab) */
size_t ifos_psi_elems(const ifos_geom *g, int n);

typedef struct ifos_fields {       /* device, device layout */
    float *vx, *vy, *vz, *sxx, *syy, *szz, *sxy, *syz, *sxz;
} ifos_fields;

typedef struct ifos_model {        /* device, device layout */
    const float *rho, *pi, *u;     /* cell values, halos [0..N+1] valid */
    const float *uipjp, *ujpkp, *uipkp, *rjp, *rkp, *rip; /* av_mat output */
} ifos_model;

typedef struct ifos_psi {          /* device, compact slab layout */
    float *p[9];
} ifos_psi;

/* kernel categories of the optional launch timer */
enum {
    IFOS_K_V_INTERIOR = 0, IFOS_K_V_CPML, IFOS_K_V_FUSED,
    IFOS_K_S_INTERIOR, IFOS_K_S_CPML, IFOS_K_S_FUSED,
    IFOS_K_DFT, IFOS_K_GRADIENT, IFOS_K_SOURCES, IFOS_K_OTHER,
    IFOS_K_VS_FUSED,  /* the batched path's fused stress + velocity step */
    IFOS_K_NCAT
};

/* Launch timer: when enabled, every kernel launch is bracketed by two HIP
 * events on its stream; ifos_prof_read synchronizes and returns, per
 * category, the summed kernel milliseconds and the launch count. */
int ifos_prof_enable(int on);
int ifos_prof_reset(void);
int ifos_prof_read(double *ms, long *count, int ncat);

const char *ifos_last_error(void);
int ifos_device_sync(void);

/* device memory and copies (synchronous), for callers without a HIP
 * runtime of their own */
int ifos_dmalloc(void **p, size_t bytes);
int ifos_dfree(void *p);
int ifos_memcpy_h2d(void *dst, const void *src, size_t bytes);
int ifos_memcpy_d2h(void *dst, const void *src, size_t bytes);
int ifos_memcpy_d2d(void *dst, const void *src, size_t bytes);
int ifos_dmemset(void *p, int value, size_t bytes);
int ifos_device_count(int *n);
int ifos_device_name(char *name, int len);

/* --- host helpers (same arithmetic as the reference, run on the CPU) --- */
/* CPML_coeff.c:87-324 -> prof[18*(2fw+1)] host array */
int ifos_cpml_profiles(const ifos_geom *g, float vppml, float fpml,
                       float *prof);
/* the alpha_prime profiles CPML_coeff also returns (not used by the
 * kernels): out[6*(2fw+1)] = x, x_half, y, y_half, z, z_half, 1-based */
int ifos_cpml_alpha(const ifos_geom *g, float fpml, float *out);
/* wavelet.c:47-90 (shapes 1, 2, 4, 5) -> sig[nt+2] indexed by nt */
int ifos_wavelet(const ifos_geom *g, int shape, float tshift, float fc,
                 float amp, float *sig);
/* ifos3d.c:745-750 */
int ifos_ntast(float fmax, int tast, float dt);

/* --- device kernels (one launch each, asynchronous on `stream`) --- */

/* update_v_ssg.c:27-130 (INTERIOR), update_v_ssg_CPML.c:30-380 (CPML). */
int ifos_update_v(const ifos_geom *g, int region, const ifos_fields *f,
                  const ifos_model *m, const float *d_prof,
                  const ifos_psi *psi_s, void *stream);
/* update_s_ssg_elastic.c:28-147 / update_s_ssg_CPML_elastic.c:30-480 */
int ifos_update_s(const ifos_geom *g, int region, const ifos_fields *f,
                  const ifos_model *m, const float *d_prof,
                  const ifos_psi *psi_v, void *stream);
/* viscoelastic arrays (device, device layout): r = rxx ryy rzz rxy ryz rxz
 * memory variables; taus, taup relaxation-time ratios (cell values, halos
 * [0..N+1] valid); tausipjp, tausjpkp, tausipkp their averages
 * (ifos_av_tau); eta = DT / relaxation time of the mechanism; L = 1 */
#define IFOS_MAX_L 4
typedef struct ifos_visco {
    float *r[6];
    const float *taus, *taup, *tausipjp, *tausjpkp, *tausipkp;
    float eta;
    int L;
    /* L > 1 (beyond the reference, which hard-codes one mechanism,
     * read_par_json.c:268-282): eta of mechanisms 2..L; each r[n] then
     * holds L fields of the device layout back to back (mechanism l at
     * r[n] + l * ifos_field_elems) */
    float eta_more[IFOS_MAX_L - 1];
} ifos_visco;
/* update_s_ssg.c:28-245 (INTERIOR) / update_s_ssg_CPML.c:30-772 (CPML): the
 * viscoelastic stress update; L = 1 is the reference's update exactly, and
 * L mechanisms (up to IFOS_MAX_L) sum their memory variables where the
 * reference uses its one, each mechanism with its own eta */
int ifos_update_s_visco(const ifos_geom *g, int region, const ifos_fields *f,
                        const ifos_model *m, const ifos_visco *visco,
                        const float *d_prof, const ifos_psi *psi_v,
                        void *stream);
/* av_mat.c:56-60: arithmetic averages of taus (halos wrapped first) */
int ifos_av_tau(const ifos_geom *g, const float *d_taus, float *d_tausipjp,
                float *d_tausjpkp, float *d_tausipkp, void *stream);
/* the same interior updates over an explicit box {i0,i1,j0,j1,k0,k1}
 * inside [1..N]^3 (update_v and update_s_elastic take the box as
 * arguments), without CPML */
int ifos_update_v_box(const ifos_geom *g, const int box[6],
                      const ifos_fields *f, const ifos_model *m, void *stream);
int ifos_update_s_box(const ifos_geom *g, const int box[6],
                      const ifos_fields *f, const ifos_model *m, void *stream);
/* body forces (update_v_ssg.c:474-510), d_src [nsrc][4] = i, j, k, type;
 * d_sig [nsrc][nt+2] by nt; alpha/beta in degrees (ALPHA, BETA) */
int ifos_add_forces(const ifos_geom *g, const ifos_fields *f,
                    const ifos_model *m, const int *d_src, const float *d_sig,
                    int nsrc, int nt, float alpha, float beta, void *stream);
/* raw adjoint traces (update_v_ssg.c:512-522), d_rec [ntr][3] = i, j, k;
 * d_adj [3][ntr][nt+2] by nt (x, y, z).  d_chain [ntr][2] from
 * ifos_rec_chain (or NULL: found on the device each call) orders the adds
 * of receivers sharing a cell as the reference does. */
int ifos_add_adjoint(const ifos_geom *g, const ifos_fields *f,
                     const int *d_rec, const int *d_chain, const float *d_adj,
                     int ntr, int nt, void *stream);
/* host: same-cell chain of a receiver list, chain [ntr][2] =
 * (first of its cell?, next receiver of the cell or -1) */
int ifos_rec_chain(const int *rec_ijk, int ntr, int *chain);
/* psource.c:26-54 (type-1 sources only) */
int ifos_psource(const ifos_geom *g, const ifos_fields *f, const int *d_src,
                 const float *d_sig, int nsrc, int nt, void *stream);
/* surface_ssg_elastic.c:26-112 (free surface at j = 1) */
int ifos_surface(const ifos_geom *g, const ifos_fields *f,
                 const ifos_model *m, void *stream);
/* seismo_ssg.c:27-53: d_sec [3][ntr][ns+1] by sample */
int ifos_seismo(const ifos_geom *g, const ifos_fields *f, const int *d_rec,
                int ntr, int ins, int ns, float *d_sec, void *stream);
/* disc_fourier.c:28-62 on [1..N]^3.  d_F[6] = Re x, y, z, Im x, y, z; each
 * nf consecutive device-layout fields.  d_C[6] = Kahan compensations (same
 * shape) for IFOS_DFT_KAHAN, else NULL.  t = nt*DT (back 0) or
 * TIME - nt*DT (back 1); trig factors computed on the host in double. */
int ifos_discfourier(const ifos_geom *g, const ifos_fields *f, float *const d_F[6],
                     float *const d_C[6], const float *finv, int nf, int nt,
                     int ntast, int back, int precision, void *stream);
/* gradient_F.c:30-170: accumulates d_grad[3] (vp, vs, rho) on [1..N]^3 */
int ifos_gradient_F(const ifos_geom *g, float *const d_Ff[6],
                    float *const d_Fb[6], const ifos_model *m,
                    const float *finv, int nf, float *const d_grad[3],
                    void *stream);
/* matcopy.c:31-307 single rank: periodic halos of a model array */
int ifos_matcopy_wrap(const ifos_geom *g, float *d_m, void *stream);
/* hess_F.c:28-192: accumulates the diagonal (pseudo-)Hessian d_hess[3]
 * (vp, vs, rho) on [1..N]^3 from the forward DFT fields d_Ff (frequency l at
 * + (l-1) * ifos_field_elems) and the receiver-side DFT fields d_Fb of
 * receivers n = 1..ntr_hess (receiver n, frequency l at
 * + (n*nf + l - 1) * ifos_field_elems, as the reference's Fbv* index them),
 * receivers outer, frequencies inner */
int ifos_hess_F(const ifos_geom *g, float *const d_Ff[6], float *const d_Fb[6], int ntr_hess,
                const ifos_model *m, const float *finv, int nf, float *const d_hess[3], void *stream);
/* readseis.c:27-66 (SU files): reads ntr traces of ns samples from path
 * into the device section d_section [ntr][ns+1] (sample j at index j,
 * index 0 set to 0), skipping each 240-byte trace header; samples past the
 * end of a short file are 0, as in readseis.c.  An unreadable file is an
 * IFOS_EINVAL error. */
int ifos_su_read(const char *path, int ntr, int ns, float *d_section);
/* outseis.c:30-275 with SEIS_FORMAT 1 (SU) as saveseis.c:38-80 calls it:
 * writes the device section d_section [ntr][ns+1] as ntr SU traces.
 * rec_xyz [ntr][3] holds each trace's receiver grid indices (x, y, z) and
 * rec_id [ntr] its global receiver number (tracl; NULL = 1..ntr); src_xyz
 * (x, y, z in metres, as srcpos[1..3]) is the shot's position, or NULL for
 * the reference's nsrc != 1 case (zero source coordinates).  Header words
 * the reference leaves unassigned (cdp, selev, swdep, unass) are written as
 * 0. */
int ifos_su_write(const char *path, const float *d_section, int ntr, int ns, const int *rec_xyz,
                  const int *rec_id, const float *src_xyz, float dx, float dy, float dz, float dt, int ndt);
/* checkfd_ssg.c:77-158 and :473-510: *cmax = the largest phase velocity at
 * infinite frequency over [1..N]^3, max of sqrt(pi (1 + L taup) / rho) and
 * sqrt(u (1 + L taus) / rho) (without the L terms when L = 0; d_taup and
 * d_taus may then be NULL), in the reference's float/double mix; *dt_stab =
 * min(DX, DY, DZ) / (gamma sqrt(3) cmax) with the 4th-order Holberg gamma
 * 1.184614.  Returns IFOS_EUNSTABLE when g->dt > *dt_stab (the reference
 * aborts there), IFOS_OK otherwise; cmax / dt_stab may be NULL. */
int ifos_check_stability(const ifos_geom *g, const float *d_rho, const float *d_pi, const float *d_u,
                         const float *d_taup, const float *d_taus, int L, float *cmax, float *dt_stab,
                         void *stream);
/* av_mat.c:28-80: writes the six averaged arrays of *out */
int ifos_av_mat(const ifos_geom *g, const float *d_rho, const float *d_u,
                float *d_uipjp, float *d_ujpkp, float *d_uipkp, float *d_rjp,
                float *d_rkp, float *d_rip, void *stream);
/* residual.c:27-56: d_obs, d_syn [3][ntr][ns+1]; d_adj [3][ntr][ns+2];
 * *l2 (host) += sum of squares, reduced in FP64 */
int ifos_residual(const ifos_geom *g, const float *d_obs, const float *d_syn,
                  float *d_adj, int ntr, int ns, double *l2, void *stream);
/* ifos_residual without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as ifos_residual, bit for bit.  The FP64 partial
 * sums live in a static device array of 3 * IFOS_RESIDUAL_NOALLOC_MAX_TRACES
 * + 1 doubles (384 KiB), one per device, held until the process ends;
 * ntr above IFOS_RESIDUAL_NOALLOC_MAX_TRACES returns IFOS_EINVAL.  Calls on
 * one device share that array: calls that overlap on one device race. */
#define IFOS_RESIDUAL_NOALLOC_MAX_TRACES 16384
int ifos_residual_noalloc(const ifos_geom *g, const float *d_obs, const float *d_syn,
                          float *d_adj, int ntr, int ns, double *l2, void *stream);
/* fill n floats with a value (graph-capture safe zeroing) */
int ifos_fill(float *d, size_t n, float value, void *stream);
/* plain streaming copy of n floats (n % 4 == 0, 16-byte aligned) */
int ifos_copy(float *d_dst, const float *d_src, size_t n, void *stream);

/* --- one gradient shot, the plain path ---------------------------------
 * Forward NT steps, residual, adjoint NT steps, gradient_F, in the call
 * order of ifos3d.c:712-975, one kernel per reference routine per step.
 * A context owns every device buffer of one shot (fields, CPML memory,
 * averaged model, DFT accumulators, traces). */
typedef struct ifos_ctx ifos_ctx;

typedef struct ifos_shot {
    int src_i, src_j, src_k, src_type;   /* grid indices, srcpos_loc[7] */
    float tshift, fc, amp;               /* srcpos_loc[4..6] */
    int source_shape;                    /* SOURCE_SHAPE */
    int ntr;
    const int *rec_ijk;                  /* host [ntr][3] */
    const float *d_obs;                  /* device [3][ntr][nt+1] or NULL */
    int want_adjoint;                    /* 0: forward only */
} ifos_shot;

int ifos_ctx_create(ifos_ctx **out, const ifos_geom *g, int nf_max,
                    int ntr_max, int dft_precision);
int ifos_ctx_destroy(ifos_ctx *c);
/* model in device layout (rho, pi, u); copied, halos wrapped, averaged */
int ifos_ctx_set_model(ifos_ctx *c, const float *d_rho, const float *d_pi,
                       const float *d_u, void *stream);
int ifos_ctx_set_cpml(ifos_ctx *c, float vppml, float fpml);
/* DFT frequencies and TAST (ntast = ifos_ntast(finv[nf-1], tast, dt)) */
int ifos_ctx_set_freqs(ifos_ctx *c, const float *finv, int nf, int tast);
/* L = 1: viscoelastic stress update with one relaxation mechanism
 * (update_s_ssg.c), taus/taup in device layout (copied, halos wrapped,
 * averaged), eta = DT / relaxation time; L = 0: elastic (the default) */
int ifos_ctx_set_visco(ifos_ctx *c, const float *d_taus, const float *d_taup,
                       float eta, int L, void *stream);
/* L = 1..IFOS_MAX_L relaxation mechanisms, eta[l] = DT / relaxation time
 * of mechanism l (the reference reads one, read_par_json.c:268-282; L > 1
 * sums the mechanisms' memory variables where update_s_ssg.c uses its one);
 * L = 0: elastic.  ifos_ctx_set_visco is this with L <= 1. */
int ifos_ctx_set_visco_n(ifos_ctx *c, const float *d_taus, const float *d_taup,
                         const float *eta, int L, void *stream);
/* runs one shot; d_syn [3][ntr][nt+1] (device, written); d_grad[3]
 * (device layout, accumulated); *l2 (host, accumulated) */
int ifos_ctx_run_shot(ifos_ctx *c, const ifos_shot *s, float *d_syn,
                      float *const d_grad[3], double *l2, void *stream);
/* Time-domain gradient with checkpoints (no counterpart in the reference,
 * whose gradient is the frequency-domain gradient_F.c): the gradient_F.c
 * products summed over every time sample of the shot instead of over DFT
 * frequencies -- lambda and mu from the displacements (velocities
 * integrated from the end of the record), rho from the velocities, each
 * sample weighted by DT, forward sample n paired with adjoint step
 * NT - 1 - n (the DFT times of disc_fourier.c:36-37) for n = 1..NT-2; the
 * lambda / mu / rho sums are formed in FP64 and added to d_grad (vp, vs, rho
 * as gradient_F.c:135-139).  The forward state is stored every `every` steps
 * (<= 0: round(sqrt(NT))) and each segment is recomputed from its
 * checkpoint during the adjoint pass; the context keeps these buffers
 * (ifos_td_bytes) until it is destroyed or `every` changes.  Needs
 * want_adjoint, d_obs and d_syn; d_syn receives the forward seismograms,
 * *l2 the misfit as ifos_ctx_run_shot. */
int ifos_ctx_run_shot_td(ifos_ctx *c, const ifos_shot *s, int every, float *d_syn,
                         float *const d_grad[3], double *l2, void *stream);
/* device bytes the time-domain gradient buffers take (L: viscoelastic) */
size_t ifos_td_bytes(const ifos_geom *g, int every, int L);
/* Consistent (exact discrete) adjoint gradient -- no counterpart in the
 * reference, whose adjoint re-runs the forward scheme with the reversed,
 * integrated residuals (residual.c:43-53, injected twice:
 * update_v_ssg.c:512-522, update_v_ssg_CPML.c:384-393) and whose gradient
 * is the frequency-domain estimate of gradient_F.c.  The misfit is
 * J = 1/2 sum over samples 1..NT, receivers and components of
 * (syn - obs)^2 of the plain path's own forward scheme (ifos_ctx_run_shot's
 * forward, bit-identical), and d_grad receives (added to it) its exact
 * derivative with respect to vp, vs and rho per cell (pi = rho vp^2,
 * u = rho vs^2, through av_mat's averages and matcopy's halo copies): the
 * transpose of every time step run backwards (hand-written adjoint
 * stencils), the model derivatives summed in FP64.  The forward state is
 * stored every `every` steps (<= 0: round(sqrt(NT))) and each segment is
 * recomputed once during the reverse pass; the context keeps these buffers
 * (ifos_adj_bytes) until it is destroyed or `every` or ntr changes.  Elastic
 * update, no free surface and an explosive source (type 1) only; needs
 * d_obs, d_syn (receives the forward seismograms) and d_grad; *J the misfit. */
int ifos_ctx_run_shot_adj(ifos_ctx *c, const ifos_shot *s, int every, float *d_syn,
                          float *const d_grad[3], double *J, void *stream);
/* device bytes the consistent-adjoint buffers take */
size_t ifos_adj_bytes(const ifos_geom *g, int every, int ntr);
/* device pointers owned by the context (fields, DFT accumulators) */
int ifos_ctx_fields(ifos_ctx *c, ifos_fields *f);
int ifos_ctx_dft(ifos_ctx *c, int back, float *d_F[6]);

/* ---- batched gradient path (no counterpart in the reference: it runs the
 * ifos3d.c:712-975 shot sequence for several shots of one model at once).
 * Per step one fused velocity launch and one fused stress launch cover all
 * shots of the batch; the DFT is accumulated in chunks of `chunk` samples
 * from a velocity snapshot ring.  Results are bit-identical to
 * ifos_ctx_run_shot on the same shots in the same order. */
typedef struct ifos_batch ifos_batch;
/* device bytes a batch of these dimensions allocates */
size_t ifos_batch_bytes(const ifos_geom *g, int nshots, int nf_max,
                        int ntr_max, int dft_precision, int chunk);
/* nshots: shots per batch; nf_max <= 8; chunk in [1, 32] */
int ifos_batch_create(ifos_batch **out, const ifos_geom *g, int nshots,
                      int nf_max, int ntr_max, int dft_precision, int chunk);
int ifos_batch_destroy(ifos_batch *b);
/* model in device layout (rho, pi, u): copied, halos wrapped, averaged */
int ifos_batch_set_model(ifos_batch *b, const float *d_rho,
                         const float *d_pi, const float *d_u);
int ifos_batch_set_cpml(ifos_batch *b, float vppml, float fpml);
int ifos_batch_set_freqs(ifos_batch *b, const float *finv, int nf, int tast);
/* as ifos_ctx_set_visco, for every shot of the batch */
int ifos_batch_set_visco(ifos_batch *b, const float *d_taus, const float *d_taup,
                         float eta, int L);
/* as ifos_ctx_set_visco_n, for every shot of the batch */
int ifos_batch_set_visco_n(ifos_batch *b, const float *d_taus, const float *d_taup,
                           const float *eta, int L);
/* runs ns (<= nshots) shots with equal ntr and want_adjoint, synchronously.
 * d_syn: device [ns][3][ntr][nt+1] or NULL; d_grad[3]: device layout,
 * accumulated over the shots in order (NULL: no gradient); l2: host [ns]
 * (per-shot misfit, FP64) or NULL; use_graph: capture the time loop once
 * into a hipGraph and replay it */
int ifos_batch_run(ifos_batch *b, const ifos_shot *shots, int ns,
                   float *d_syn, float *const d_grad[3], double *l2,
                   int use_graph);
/* grid levels of uneven size (no counterpart in the reference, whose
 * coarse-to-fine levels run one after the other): forward solves of nlev
 * batch contexts, each on its own grid, run together.  Level l runs ns[l]
 * shots shots[l] (want_adjoint = 0) and, when d_syn and d_syn[l] are not
 * NULL, writes their seismograms to d_syn[l] [ns[l]][3][ntr][nt+1] as
 * ifos_batch_run on that context alone would, bit for bit.  The levels'
 * time loops are enqueued on the contexts' own streams (use_graph = 0), or
 * captured once into one graph with one branch per level and replayed
 * (use_graph = 1; the graph is kept by lv[0] until a context's launch
 * state changes).  Synchronous; every context appears once. */
#define IFOS_MAX_LEVELS 8
int ifos_levels_run(ifos_batch *const *lv, int nlev, const ifos_shot *const *shots, const int *ns,
                    float *const *d_syn, int use_graph);
/* tile shapes of the velocity and stress march launches, each
 * {TI, jc, group, split, pipe, kb, sb, lds}: TI wavefronts per workgroup (1, 2, 4, 8);
 * jc j-planes per tile (<= 0: automatic); group: tiles of one shot before
 * the next shot (1: shots fastest, <= 0: one shot after the other); split:
 * 1 = interior box and CPML slabs in separate launches, 0 = one CPML-aware
 * launch; pipe: 1 = the next plane's operands are loaded one plane ahead;
 * kb: the workgroup's wavefronts stand kb side by side along k (64 k each);
 * sb: sb shots of one tile share a workgroup (and its model loads); the
 * remaining TI / (kb sb) wavefronts go along i (kb sb divides TI); lds: 1 =
 * the in-plane neighbours are staged in LDS (used where every row is 64 k
 * wide and kb = sb = 1, TI >= 2; elsewhere ignored).  NULL leaves a shape
 * unchanged.  A new batch starts with {4, 32, 0, 0, 0, 1, 1, 1} for both. */
#define IFOS_TILE_FIELDS 8
int ifos_batch_set_tiles(ifos_batch *b, const int cfg_v[IFOS_TILE_FIELDS], const int cfg_s[IFOS_TILE_FIELDS]);
/* the tile shape (same fields; group, split and sb have no effect) of the
 * one-shot march launches: ifos_update_v / ifos_update_s(_visco) on the
 * interior or the whole domain, the shot contexts and the drop-in.
 * Process-wide, default {4, 32, 0, 0, 0, 1, 1, 1}; set it between launches,
 * not while another thread launches. */
int ifos_set_plain_tiles(const int cfg[IFOS_TILE_FIELDS]);
int ifos_get_plain_tiles(int cfg[IFOS_TILE_FIELDS]);
/* the autotuner's candidate shapes (IFOS_TILE_FIELDS ints each, at most
 * max); returns their number */
int ifos_batch_tune_candidates(int *cfg, int max);
/* times every candidate on the batch's buffers for ns shots (reps launches
 * each), keeps the fastest per launch kind, returns their indices and the
 * per-candidate average ms (arrays of the candidates' length, or NULL);
 * clears the wavefields */
int ifos_batch_autotune(ifos_batch *b, int ns, int reps, int *best_v,
                        int *best_s, float *ms_v, float *ms_s);
/* the fused step (stress of step n and velocity of step n+1 in one launch,
 * double-buffered fields): off by default, used when on and there is no free
 * surface, no source or receiver inside a CPML slab and L = 0; ti = 8 or 16
 * workgroup rows, jc j-planes per tile (<= 0: automatic) */
int ifos_batch_set_fused(ifos_batch *b, int on, int ti, int jc);
/* the fused-step timings of the last ifos_batch_autotune (which also picks
 * the fused step's ti and jc, and turns the fused step on only if it beats
 * the best velocity + stress launch pair): out[3c] = ti, jc, average ms;
 * returns the count (at most max) */
int ifos_batch_tune_fused(const ifos_batch *b, float *out, int max);
/* 1 if the last ifos_batch_run used the fused step */
int ifos_batch_last_fused(const ifos_batch *b);
/* DFT accumulators of one shot of the last batch: frequency l of
 * component n (Re x y z, Im x y z) at d_F[n] + l * ifos_field_elems(g) */
int ifos_batch_dft(ifos_batch *b, int shot, int back, float *d_F[6]);

/* ---- optimiser steps on the device, dot products reduced in FP64 (the
 * reference accumulates them in FP32 and all-reduces MPI_FLOAT) ---- */
typedef struct ifos_lbfgs ifos_lbfgs;
/* the L-BFGS history of ifos3d.c:373-381: bfgsgrad and bfgsmod, bfgsnum
 * slots of numpar * NX*NY*NZ floats (numpar 1: grad1 only, 2: grad1 and
 * grad2, lbfgs.c:49-71), and bfgsscale; all zero */
int ifos_lbfgs_create(const ifos_geom *g, int numpar, int bfgsnum, ifos_lbfgs **out);
void ifos_lbfgs_destroy(ifos_lbfgs *b);
/* device pointers of slot w (1..bfgsnum) of bfgsmod (the model step that
 * modelupdate.c:146-149 writes, cell ((j-1)*NX + i-1)*NZ + k-1, grad2's
 * block after grad1's) and of bfgsgrad */
float *ifos_lbfgs_mod(ifos_lbfgs *b, int w);
float *ifos_lbfgs_grad(ifos_lbfgs *b, int w);
/* bfgsscale[0..bfgsnum-1] (1 / y.s of each slot) */
int ifos_lbfgs_scale(const ifos_lbfgs *b, double *scale);
/* lbfgs_save.c:25-54: slot 1 of bfgsgrad = the gradient (first iteration) */
int ifos_lbfgs_savegrad(ifos_lbfgs *b, const float *d_grad1, const float *d_grad2);
/* lbfgs.c:26-181 for iteration >= 2: y = previous - current gradient,
 * rho = 1 / y.s, h0 = y.s / y.y, the two-loop recursion over the last
 * bfgsnum pairs, the current gradient saved, d_grad1 (and d_grad2) replaced
 * by the update direction.  *h0_out (may be NULL) receives h0. */
int ifos_lbfgs_update(ifos_lbfgs *b, float *d_grad1, float *d_grad2, int iteration, double *h0_out);
/* conjugategrad.c:26-110 on the device-layout fields: cdf 1 copies the
 * gradients to all six priors; else beta from the six sums (written to
 * sums[6] when not NULL), the reference's clamps, the new direction in
 * d_grad and d_prior[3..5]; beta[3] receives the factors */
int ifos_conjugate(const ifos_geom *g, float *const d_grad[3], float *const d_prior[6], float beta[3], int cdf,
                   double *sums);
/* precongrad.c:28-336: tapers the three gradients (device layout, [1..N]^3)
 * for DAMPTYPE 0..3: around the sources (src_xyz[nsrc][3], grid x y z, i.e.
 * iround(srcpos / DX ...)), around the receivers (rec_xyz[ntr][3], recpos
 * order x y z) and towards the model edges; nglob = NXG NYG NZG (NULL: this
 * grid), pos = POS[1..3] (NULL: 0).  Keeps the reference's quirks: the
 * global indices are 0 outside DAMPTYPE 1 and 2, and FREE_SURF zeroes the
 * y factor. */
int ifos_precon_grad(const ifos_geom *g, float *const d_grad[3], int damptype, const int *src_xyz, int nsrc,
                     const int *rec_xyz, int ntr, const int nglob[3], const int pos[3], void *stream);
/* modelupdate.c:26-160: vp, vs, rho += step * WEIGHT * grad * (VP0, VS0,
 * RHO0), rho, pi, u rewritten; without L-BFGS the gradients are first divided
 * by their maxima; with it, slot d_bfgsmod (compact, ifos_lbfgs_mod) receives
 * the normalised vp (and, numpar 2, vs) steps.  maxima[3] (may be NULL):
 * vp, vs, rho maxima of the model before the update. */
int ifos_modelupdate(const ifos_geom *g, float *const d_grad[3], float *d_rho, float *d_pi, float *d_u,
                     float *d_bfgsmod, int numpar, int lbfgs, float step, const float weight[3], float vp0,
                     float vs0, float rho0, float maxima[3], void *stream);
/* steplength.c:26-91 (host): the parabola through (step[0..2], L2[0..2]);
 * step[3] and L2[3] receive its clamped minimum, step[4] the next test
 * step (TESTSTEP = teststep) */
int ifos_steplength(float L2[4], float step[5], int lbfgs, int it_group, float teststep);
/* the same parabola fit in FP64 on the device (no counterpart in the
 * reference, whose steplength.c:26-91 is float): d_L2[4] and d_step[5] are
 * device doubles, typically the FP64 misfits of the test steps; one thread on
 * `stream`, no host round trip */
int ifos_steplength_f64(double *d_L2, double *d_step, int lbfgs, int it_group, double teststep, void *stream);

#ifdef __cplusplus
}
#endif
#endif
