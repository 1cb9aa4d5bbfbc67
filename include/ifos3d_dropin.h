/*
 * ifos3d_dropin.h -- the reference-compatible entry points of
 * libifos3d_dropin.so: the IFOS3D hot-path routines with their own
 * prototypes (src/fd.h of IFOS3D), executed on an MI355X.
 *
 * Link the IFOS3D driver (ifos3d.c and the routines it keeps) against
 * libifos3d_dropin.so instead of the object files of the routines below;
 * nothing in the caller changes.  The library reads the caller's globals
 * (it does not define them): DT DX DY DZ TIME NT NX NY NZ FDORDER FDCOEFF
 * ABS_TYPE FW FREE_SURF L POS NPROCX NPROCY NPROCZ SEISMO ALPHA BETA VPPML
 * FPML (globvar.h:27-62 of IFOS3D).
 *
 * Residency: every NR array a routine receives is mirrored on the device the
 * first time it is seen (keyed by its float*** pointer) and stays there.
 * Host copies are refreshed only where the caller reads them:
 *   - seismo() writes the host section arrays at every call;
 *   - the NaN probe cell vy[NY/2][NX/2][NZ/2] (ifos3d.c:761) is copied back
 *     after every velocity update;
 *   - av_mat() uploads rho/pi/u (after the caller's matcopy) and writes the
 *     averaged arrays back;
 *   - gradient_F() uploads grad1..3, accumulates on the device and writes
 *     them back;
 *   - sources, receiver positions and traces are uploaded at nt == 1 of
 *     every propagation (the caller rewrites them per shot).
 * Other host wavefield/DFT arrays are stale while the device works.
 *
 * Supported configuration: one rank (NPROCX*NPROCY*NPROCZ == 1), FDORDER 4,
 * ABS_TYPE 1 (CPML), L 0 (elastic) or 1 (one relaxation mechanism: update_s,
 * update_s_CPML), SEISMO 1.  Anything else stops the run
 * with a message, as the reference's err() does (util.c:30-40).
 */
#ifndef IFOS3D_DROPIN_H
#define IFOS3D_DROPIN_H

#ifdef __cplusplus
extern "C" {
#endif

/* fd.h:281-285 (update_v_ssg.c:27) */
double update_v(int nx1, int nx2, int ny1, int ny2, int nz1, int nz2, int nt,
                float ***vx, float ***vy, float ***vz, float ***sxx,
                float ***syy, float ***szz, float ***sxy, float ***syz,
                float ***sxz, float ***rho, float ***rjp, float ***rkp,
                float ***rip, float **srcpos_loc, float **signals,
                float **signaly, float **signalz, int nsrc,
                float ***absorb_coeff, int back);

/* fd.h:287-296 (update_v_ssg_CPML.c:30) */
double update_v_CPML(int nx1, int nx2, int ny1, int ny2, int nz1, int nz2,
                     int nt, float ***vx, float ***vy, float ***vz,
                     float ***sxx, float ***syy, float ***szz, float ***sxy,
                     float ***syz, float ***sxz, float ***rho, float ***rjp,
                     float ***rkp, float ***rip, float **srcpos_loc,
                     float **signals, float **signaly, float **signalz,
                     int nsrc, float ***absorb_coeff, int back, float *K_x,
                     float *a_x, float *b_x, float *K_x_half, float *a_x_half,
                     float *b_x_half, float *K_y, float *a_y, float *b_y,
                     float *K_y_half, float *a_y_half, float *b_y_half,
                     float *K_z, float *a_z, float *b_z, float *K_z_half,
                     float *a_z_half, float *b_z_half, float ***psi_sxx_x,
                     float ***psi_sxy_x, float ***psi_sxz_x,
                     float ***psi_sxy_y, float ***psi_syy_y,
                     float ***psi_syz_y, float ***psi_sxz_z,
                     float ***psi_syz_z, float ***psi_szz_z);

/* fd.h:402 (update_s_ssg_elastic.c:28) */
double update_s_elastic(int nx1, int nx2, int ny1, int ny2, int nz1, int nz2,
                        float ***vx, float ***vy, float ***vz, float ***sxx,
                        float ***syy, float ***szz, float ***sxy,
                        float ***syz, float ***sxz, float ***pi, float ***u,
                        float ***uipjp, float ***ujpkp, float ***uipkp);

/* fd.h:404 (update_s_ssg_CPML_elastic.c:30) */
double update_s_CPML_elastic(int nx1, int nx2, int ny1, int ny2, int nz1,
                             int nz2, float ***vx, float ***vy, float ***vz,
                             float ***sxx, float ***syy, float ***szz,
                             float ***sxy, float ***syz, float ***sxz,
                             float ***pi, float ***u, float ***uipjp,
                             float ***ujpkp, float ***uipkp, float *K_x,
                             float *a_x, float *b_x, float *K_x_half,
                             float *a_x_half, float *b_x_half, float *K_y,
                             float *a_y, float *b_y, float *K_y_half,
                             float *a_y_half, float *b_y_half, float *K_z,
                             float *a_z, float *b_z, float *K_z_half,
                             float *a_z_half, float *b_z_half,
                             float ***psi_vxx, float ***psi_vyx,
                             float ***psi_vzx, float ***psi_vxy,
                             float ***psi_vyy, float ***psi_vzy,
                             float ***psi_vxz, float ***psi_vyz,
                             float ***psi_vzz);

/* fd.h:247-254 (update_s_ssg.c:28): viscoelastic, one relaxation
 * mechanism (L = 1); the box must be the CPML_ini_elastic interior box */
double update_s(int nx1, int nx2, int ny1, int ny2, int nz1, int nz2,
                float ***vx, float ***vy, float ***vz, float ***sxx,
                float ***syy, float ***szz, float ***sxy, float ***syz,
                float ***sxz, float ***rxx, float ***ryy, float ***rzz,
                float ***rxy, float ***ryz, float ***rxz, float ***pi,
                float ***u, float ***uipjp, float ***ujpkp, float ***uipkp,
                float ***taus, float ***tausipjp, float ***tausjpkp,
                float ***tausipkp, float ***taup, float *eta);

/* fd.h:256-263 (update_s_ssg_CPML.c:30) */
double update_s_CPML(int nx1, int nx2, int ny1, int ny2, int nz1, int nz2,
                     float ***vx, float ***vy, float ***vz, float ***sxx,
                     float ***syy, float ***szz, float ***sxy, float ***syz,
                     float ***sxz, float ***rxx, float ***ryy, float ***rzz,
                     float ***rxy, float ***ryz, float ***rxz, float ***pi,
                     float ***u, float ***uipjp, float ***ujpkp,
                     float ***uipkp, float ***taus, float ***tausipjp,
                     float ***tausjpkp, float ***tausipkp, float ***taup,
                     float *eta, float *K_x, float *a_x, float *b_x,
                     float *K_x_half, float *a_x_half, float *b_x_half,
                     float *K_y, float *a_y, float *b_y, float *K_y_half,
                     float *a_y_half, float *b_y_half, float *K_z, float *a_z,
                     float *b_z, float *K_z_half, float *a_z_half,
                     float *b_z_half, float ***psi_vxx, float ***psi_vyx,
                     float ***psi_vzx, float ***psi_vxy, float ***psi_vyy,
                     float ***psi_vzy, float ***psi_vxz, float ***psi_vyz,
                     float ***psi_vzz);

/* fd.h:130 (psource.c:26) */
void psource(int nt, float ***sxx, float ***syy, float ***szz,
             float **srcpos_loc, float **signals, int nsrc);

/* fd.h:198-201 (seismo_ssg.c:27) */
void seismo(int lsamp, int ntr, int **recpos, float **sectionvx,
            float **sectionvy, float **sectionvz, float **sectiondiv,
            float **sectioncurl, float **sectionp, float ***vx, float ***vy,
            float ***vz, float ***sxx, float ***syy, float ***szz,
            float ***pi, float ***u);

/* fd.h:405 (surface_ssg_elastic.c:26) */
void surface_elastic(int ndepth, float ***u, float ***pi, float ***sxx,
                     float ***syy, float ***szz, float ***sxy, float ***syz,
                     float ***vx, float ***vy, float ***vz);

/* fd.h:408 (disc_fourier.c:28) */
void discfourier(int nx1, int nx2, int ny1, int ny2, int nz1, int nz2, int nt,
                 float ***vx, float ***vy, float ***vz, float ****Fvx,
                 float ****Fvy, float ****Fvz, float ****Fvxi, float ****Fvyi,
                 float ****Fvzi, float *finv, int nf, int ntast, int pshot,
                 int back);

/* fd.h:390 (gradient_F.c:30) */
void gradient_F(int nx, int ny, int nz, float ****fvx, float ****fvy,
                float ****fvz, float ****fivx, float ****fivy,
                float ****fivz, float ****bvx, float ****bvy, float ****bvz,
                float ****bivx, float ****bivy, float ****bivz,
                float ***gradlam, float ***gradmu, float ***gradrho, int nt,
                float ***rho, float ***pi, float ***u, float *finv, int nf,
                int iteration);

/* fd.h:64-67 (av_mat.c:28) */
void av_mat(float ***rho, float ***pi, float ***u, float ***taus,
            float ***taup, float ***uipjp, float ***ujpkp, float ***uipkp,
            float ***tausipjp, float ***tausjpkp, float ***tausipkp,
            float ***rjp, float ***rkp, float ***rip);

/* fd.h:57-62 (CPML_coeff.c:87) */
void CPML_coeff(float *K_x, float *alpha_prime_x, float *a_x, float *b_x,
                float *K_x_half, float *alpha_prime_x_half, float *a_x_half,
                float *b_x_half, float *K_y, float *alpha_prime_y, float *a_y,
                float *b_y, float *K_y_half, float *alpha_prime_y_half,
                float *a_y_half, float *b_y_half, float *K_z,
                float *alpha_prime_z, float *a_z, float *b_z,
                float *K_z_half, float *alpha_prime_z_half, float *a_z_half,
                float *b_z_half);

/* fd.h:128 (CPML_ini_elastic.c:26) */
void CPML_ini_elastic(int *xb, int *yb, int *zb);

/* fd.h:388 (zero_wavefield.c:25) */
void zero_wavefield(int NX, int NY, int NZ, float ***vx, float ***vy,
                    float ***vz, float ***sxx, float ***syy, float ***szz,
                    float ***sxy, float ***syz, float ***sxz, float ***rxx,
                    float ***ryy, float ***rzz, float ***rxy, float ***ryz,
                    float ***rxz, float ***psi_sxx_x, float ***psi_sxy_x,
                    float ***psi_sxz_x, float ***psi_sxy_y,
                    float ***psi_syy_y, float ***psi_syz_y,
                    float ***psi_sxz_z, float ***psi_syz_z,
                    float ***psi_szz_z, float ***psi_vxx, float ***psi_vyx,
                    float ***psi_vzx, float ***psi_vxy, float ***psi_vyy,
                    float ***psi_vzy, float ***psi_vxz, float ***psi_vyz,
                    float ***psi_vzz);

/* fd.h:418 (hess_F.c:28): the receiver-side DFT fields of receiver n,
 * frequency l are entry n*nf + l of fb* / fib* (written by discfourier with
 * pshot = n) */
void hess_F(int nx, int ny, int nz, float ****fvx, float ****fvy,
            float ****fvz, float ****fivx, float ****fivy, float ****fivz,
            float ****bvx, float ****bvy, float ****bvz, float ****bivx,
            float ****bivy, float ****bivz, float ***hess1, float ***hess2,
            float ***hess3, int nt, float ***rho, float ***pi, float ***u,
            float *finv, int nf, int ntr_hess);

/* fd.h:393 (zero_invers.c:26) */
void zero_invers(int NX, int NY, int NZ, float ****Ffvx, float ****Ffvy,
                 float ****Ffvz, float ****Ffivx, float ****Ffivy,
                 float ****Ffivz, float ****Fbvx, float ****Fbvy,
                 float ****Fbvz, float ****Fbivx, float ****Fbivy,
                 float ****Fbivz, int nfmax, int ntr_hess);

/* fd.h:398 (zero_grad.c:26) */
void zero_grad(int NX, int NY, int NZ, float ***grad1, float ***grad2,
               float ***grad3);

/* Optimiser steps.  They read the globals LBFGS, NUMPAR, BFGSNUM, VP0, VS0,
 * RHO0 and WEIGHT[3] (globvar.h:60-62), which the caller defines.  Dot
 * products are reduced in FP64.  bfgsgrad and bfgsscale's history lives on
 * the device, keyed by the bfgsgrad matrix: the host's bfgsgrad is not
 * written; bfgsscale[1..BFGSNUM] is. */
/* fd.h:425 (lbfgs_save.c:25) */
void lbfgs_savegrad(float ***grad1, float ***grad2, float ***grad3,
                    float **bfgsgrad1);
/* fd.h:424 (lbfgs.c:26): bfgsmod[1..BFGSNUM] (modelupdate's steps) is read
 * from the host */
void lbfgs(float ***grad1, float ***grad2, float ***grad3, float *bfgsscale,
           float **bfgsmod, float **bfgsgrad, int iteration);
/* fd.h:391 (modelupdate.c:26) */
void modelupdate(int nx, int ny, int nz, float ***gradvp, float ***gradvs,
                 float ***gradrho, float ***rho, float ***pi, float ***u,
                 float **bfgsmod1, float step, float *beta, int it_group);
/* fd.h:400 (conjugategrad.c:26) */
void conjugate(int nx, int ny, int nz, float ***grad1, float ***grad2,
               float ***grad3, float ***gradprior1, float ***gradprior2,
               float ***gradprior3, float ***gradprior4, float ***gradprior5,
               float ***gradprior6, float *beta, int iteration, int cdf);

/* --- extensions (no counterpart in the reference) --- */
/* message of the last fatal error (the routines above stop the process) */
const char *ifos_dropin_last_error(void);
/* wait for the device and release every device mirror */
void ifos_dropin_release(void);

#ifdef __cplusplus
}
#endif
#endif
