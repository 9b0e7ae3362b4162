/*
 * chapsim_gpu.h -- C ABI of the MI355X (gfx950) implementation of CHAPSim's
 * per-RK3-substep solver core.
 *
 * Conventions
 *  - Every entry point returns int: CHS_OK (0) or a CHS_E_* code; the message
 *    of the last failure on the calling thread is returned by chs_last_error().
 *    The reference has no status codes: its Fortran routines STOP
 *    (src/WRTHDL.f90:18-29).  The Fortran shim (fortran/chapsim_gpu.f90) turns
 *    a non-zero status into a message plus ERROR STOP.
 *  - Arrays keep the reference's Fortran layouts (column-major, x fastest):
 *        Q_tg(NCL1, 0:NCL2+1, NCL3, 3)   PR_tg, DPH_tg (NCL1, 0:NCL2+1, NCL3)
 *        RHSLLPHI_tg, RHS_tg (NCL1, NCL2, NCL3)   CONVH0_tg (NCL1, NCL2, NCL3, 3)
 *  - "dev" pointers are device (HBM) addresses; "host" pointers are host memory.
 *  - All work of a handle is ordered on the HIP stream given at creation.
 *    Calls on one handle must not overlap (not re-entrant per handle).
 *  - Only the wall-bounded channel (ICASE = 1) on one GPU is supported.
 */
#ifndef CHAPSIM_GPU_H
#define CHAPSIM_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CHS_OK = 0,
    CHS_E_ARG = 1,      /* invalid argument or unsupported shape              */
    CHS_E_HIP = 2,      /* HIP runtime error                                  */
    CHS_E_NOMEM = 3,    /* device allocation failed                           */
    CHS_E_RANGE = 4,    /* value outside a table range (reference: STOP)      */
    CHS_E_NOTCONV = 5   /* iterative refinement did not reach its tolerance   */
};

enum { CHS_FP64 = 0, CHS_MIXED = 1 };

typedef struct chs_handle chs_handle;

/* Mesh and coefficients, as the reference's set-up routines leave them in
 * its modules (all arrays are host memory, copied at creation):
 *   dyfi[M]  DYFI(1:NCL2), dyci[M+1] DYCI(1:NND2)   (src/COORDJR.f90:112-131)
 *   wff/wfb[M+1]  YCL2ND_WFF/WFB(1:NND2)            (src/COORDJR.f90:427-470)
 *   amph/acph/apph[M]                               (src/TDMA_COEF.f90:45-64)
 *   amvr/acvr/apvr[3*(M+1)], element (J, IDR) at (IDR-1)*(M+1) + J-1
 *                                                   (src/TDMA_COEF.f90:67-116)
 *   ak1[L], ak3[N]  modified wavenumbers            (src/possion3d_FFT99.f90:152-179)
 * Scalars DXI, DZI, DXQI, DZQI, CVISC=1/REN, VOLM_tg (src/CONSPARA.f90, src/COORDJR.f90:385-396). */
typedef struct chs_grid_desc {
    int ncl1, ncl2, ncl3;       /* L, M, N; L and N even with factors 2, 3, 5 */
    int icase;                  /* 1 = channel                                */
    int flowdrvtp;              /* FLOWDRVTP: 1 constant mass flux, 2 constant
                                   pressure gradient -CFGV/2 (cfgv below)     */
    int precision;              /* CHS_FP64 or CHS_MIXED                       */
    double dxi, dzi, dxqi, dzqi, cvisc, volm;
    const double *dyfi, *dyci, *wff, *wfb, *amph, *acph, *apph;
    const double *amvr, *acvr, *apvr, *ak1, *ak3;
    int nflow;                  /* NFLOW, the streamwise direction: 1 (x; 0 is
                                   taken as 1) or 3 (z)                        */
    double cfgv;                /* CFGV of FLOWDRVTP = 2 (READINI.f90:269)     */
} chs_grid_desc;

/* Create a solver handle: allocates the device-resident mirrors of the
 * reference's tg state (Q_tg, PR_tg, CONVH0_tg) and work arrays.
 * Replaces the allocation in MEM_ALLOCAT (src/MEMALLOCT.f90:99-106) and the
 * set-up FFT99_POIS3D_INIT (src/possion3d_FFT99.f90:54-117).
 * stream: a hipStream_t, or NULL to create one. */
int chs_create(const chs_grid_desc *desc, void *stream, chs_handle **out);
int chs_destroy(chs_handle *h);

/* Copy the tg state between host arrays (reference layouts) and the device. */
int chs_upload_state(chs_handle *h, const double *q_host, const double *pr_host, const double *convh0_host);
int chs_download_state(chs_handle *h, double *q_host, double *pr_host, double *convh0_host);
/* Device addresses of the handle's arrays (any output pointer may be NULL).
 * The mixed-precision form keeps CONVH0 in FP32 between substeps; asking for
 * convh0 writes it back to this FP64 array first, which then holds CONVH0 as
 * of this call (a later substep of that form does not update it). */
int chs_device_arrays(chs_handle *h, double **q, double **pr, double **convh0, double **dph, double **rhsllphi);

/* One RK3 substep NS (0..3) of the tg channel with time step dt, on the
 * device-resident state: SOLVERRK3_MOM_tg(NS) (src/SOLVERRK3_MOM_tg.f90:18-61):
 * convection, RHS + the flow drive of component NFLOW (constant mass flux or
 * constant pressure gradient, src/RHS_CvLpGpS_tg.f90:178-213), CN factorised x/z/y solves,
 * wall BCs, divergence, Poisson solve, projection and pressure update.
 * Asynchronous on h's stream.  With CHS_MIXED precision and NS >= 1 the
 * Poisson solve is one FP32 pass and one refinement whose residual is checked
 * on the device (alpha dt max|f - L phi| <= 0.5e-10, no NaN); the verdict is
 * read without a wait by a later substep when it is ready, or by
 * chs_divgck, chs_download_state or chs_poisson_stats, which then return
 * CHS_E_NOTCONV.  A failed check is kept on the device until a verdict is
 * read, so none is lost while the host runs ahead, and is reported once. */
int chs_rk3_substep(chs_handle *h, int ns, double dt);

/* Change CVISC = 1/REN of the handle (used by the tg substep's viscous terms
 * and CN solves and by the io momentum substep).  The reference recomputes
 * CVISC every step in BCAST_COMM_STEP (src/SOLVE.f90:167) after switching REN
 * from REINI to REN once phyTIME reaches TLGRE (src/SOLVE.f90:97-100); a
 * driver calls this at that switch.  cvisc > 0 and finite. */
int chs_set_cvisc(chs_handle *h, double cvisc);

/* Pressure Poisson solve FFT99_POIS3D_periodicxz (src/possion3d_FFT99.f90:254-488):
 *   f_dev:   RHSLLPHI(NCL1, NCL2, NCL3)          (device)
 *   dph_dev: DPH(NCL1, 0:NCL2+1, NCL3)           (device); interior written,
 *            ghost planes set to zero as the reference does (:447).
 * With CHS_MIXED precision the FFTs run in FP32 and the solve is iteratively
 * refined with an FP64 residual until max|L phi - f| <= tol * max|f| (at least
 * one refinement when the iteration cap allows any); CHS_E_NOTCONV when the
 * cap is reached first or the residual is NaN. */
int chs_poisson_fft99(chs_handle *h, const double *f_dev, double *dph_dev);
/* Tolerance and iteration cap of the mixed-precision refinement; the last
 * call's iteration count and relative residual are reported by chs_poisson_stats. */
int chs_set_refinement(chs_handle *h, double rel_tol, int max_iter);
int chs_poisson_stats(chs_handle *h, int *iters, double *rel_resid);

/* Crank-Nicolson line solves of MOMFA_tg (src/MOMFA_tg.f90:17-240) inside
 * chs_rk3_substep.  mode 0 (default): where every row of the inverse of the
 * x, z and y CN systems drops below 1e-17 of its absolute sum within 16 rows,
 * the solves run as the truncated inverses (short stencils, computed exactly
 * on the host in extended precision); otherwise, and always with mode 1, as
 * segmented Thomas solves.  Mode 0 evaluates the right-hand sides (here, in
 * chs_solverrk3_mom_io and chs_energy_substep) marching along y where each
 * block's chunk has at least 8 rows, one y row per block otherwise (small
 * grids); mode 2 always one row per block, mode 3 always marching where the
 * marching kernels fit (same results to rounding). */
int chs_set_cn_solver(chs_handle *h, int mode);
/* Which path substep NS with time step dt takes (1 = stencils, 0 = Thomas)
 * and the stencil half-widths in x, z, y (> 16 when too wide). */
int chs_cn_info(chs_handle *h, int ns, double dt, int *path, int *kx, int *kz, int *ky);

/* max|div u| of the device state: DIVGCK_tg (src/DIVGCK.f90:16-58). */
int chs_divgck(chs_handle *h, double *maxdiv_host);
/* max convective CFL rate: CFL_tg (src/CFL.f90:21-59); dt_cfl = CFL/rate. */
int chs_cfl(chs_handle *h, double *rate_host);
/* xz-plane means of the device state, the first and second moments that
 * PP_MEAN_ZX_FLOW_TG accumulates (src/processing/PP_MEAN_ZX.f90:1-237): with
 * cell-centred sums U1 = u(i)+u(i+1), U2 = v(j)+v(j+1), U3 = w(k)+w(k+1),
 * out[13 j + m] for row j = 0..NCL2-1 holds
 *   m = 0..2   mean(U_m) / 2            (U1xzL_tg(J,1:3))
 *   m = 3      mean(p)                  (U1xzL_tg(J,4))
 *   m = 4..9   mean(U_a U_b) / 4, (a,b) = 11 12 13 22 23 33   (U2xzL_tg)
 *   m = 10..12 mean(U_m p) / 2          (UPxzL_tg)
 * out is host memory of 13*NCL2 doubles.  One block per row sums in a fixed
 * order: the result is deterministic. */
int chs_mean_xz(chs_handle *h, double *out_host);
/* Time means of the xz-plane moments: PP_MEAN_T_TG (domain 0) and
 * PP_MEAN_T_Xperiodic_IO (domain 1) (src/processing/PP_MEAN_T.f90:322-370,
 * 436-470).  Takes the moments of the last chs_mean_xz (domain 0) or
 * chs_mean_xz_io (domain 1) call, which stay on the device, into the time
 * means: NSTATIS += 1, then mean = COE3 mean + COE2 xz with
 * (COE2, COE3) = (1/NSTATIS, (NSTATIS-1)/NSTATIS) for tp = 1
 * (T_Asymptotic_Average) or (1/pp_instn_sz, 1) for tp = 2 (T_Summing_average).
 * *nstatis (may be NULL) receives NSTATIS; out_host (may be NULL) the time
 * means, 13*NCL2 (domain 0) or 37*NCL2 (domain 1) doubles in the column order
 * of the xz means. */
int chs_mean_t(chs_handle *h, int domain, int tp, int pp_instn_sz, int *nstatis, double *out_host);
/* PP_TMEAN_INI_ZERO_TG / _IO (src/processing/PP_MEAN_T.f90:29-54, 110-319):
 * NSTATIS = 0 and the time means of the domain set to zero. */
int chs_mean_t_reset(chs_handle *h, int domain);
/* DPGRNS of the last substep (the flow-drive term: the constant-mass-flux
 * correction, or -CFGV/2 TALP DT for FLOWDRVTP = 2). */
int chs_last_dpgrns(chs_handle *h, double *dpgrns_host);

/* Batched tridiagonal line solves on device arrays, the four reference solvers:
 *   cyclic = 0: TDMAIJJ_nonCYC / TDMAIJI_nonCYC (src/TDMAIJJ_nonCYC.f90:67-123,
 *               src/TDMAIJI_nonCYC.f90:67-123); bc_dev may be NULL (zero) or
 *               hold BC(line, 1:2) column-major: bc[l], bc[nlines + l].
 *   cyclic = 1: TDMAIJJ_CYC / TDMAIJI_CYC (src/TDMAIJJ_CYC.f90:67-157,
 *               src/TDMAIJI_CYC.f90:67-157); bc_dev ignored.
 * Element (line l, row j) of a, b, c, r lives at l*line_stride + j*elem_stride:
 *   IJJ form (march along index 2):  elem_stride = ISZ, line_stride = 1
 *   IJI form (march along index 1):  elem_stride = 1,   line_stride = ISZ
 * r is overwritten by the solution.  n >= 3.  stream may be NULL (default). */
int chs_tdma_batched(const double *a_dev, const double *b_dev, const double *c_dev, double *r_dev,
                     const double *bc_dev, int n, int nlines, long long elem_stride,
                     long long line_stride, int cyclic, void *stream);

/* chs_tdma_batched without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Lines the segmented
 * kernel splits need no workspace; the others (e.g. cyclic lines of odd
 * length, or more than 512 rows) take 3*n*nlines doubles from a static array
 * of 4194304 doubles (32 MiB) that each GPU holds until the process ends, and
 * larger such calls return CHS_E_ARG.  Calls that overlap on one GPU share that
 * array and race; its contents are not cleared between calls (nothing in it
 * is read before it is written).  Same arguments and results as
 * chs_tdma_batched. */
int chs_tdma_batched_noalloc(const double *a_dev, const double *b_dev, const double *c_dev, double *r_dev,
                             const double *bc_dev, int n, int nlines, long long elem_stride,
                             long long line_stride, int cyclic, void *stream);

/* The same on host arrays (what the Fortran subroutines receive): copies the
 * arrays to the device, solves, copies r back; synchronous.  For a per-call
 * drop-in of TDMAIJ?_(non)CYC on host data; PCIe-bound. */
int chs_tdma_batched_host(const double *a, const double *b, const double *c, double *r,
                          const double *bc, int n, int nlines, long long elem_stride,
                          long long line_stride, int cyclic);

/* ---- thermophysical properties (THERM_PROP_UPDATE) ---------------------- */
typedef struct chs_props chs_props;
enum { CHS_PROPS_TABLE = 0, CHS_PROPS_SURROGATE = 1 };

/* Property table after NIST_TABLE_SCALING (src/Flow_State.f90:343-440): n
 * knots (4 <= n <= 4096) sorted by H, columns H, T, D, M, K, CP (host arrays;
 * H and T strictly increasing).  Builds the cubic splines H -> T and
 * T -> D, M, K, CP (NIST_SPLINE_H / NIST_SPLINE_T with CUBIC_SPLINE,
 * src/Flow_State.f90:916-1014, 1693-1788) and uploads them.
 * stream: a hipStream_t, or NULL to create one. */
int chs_props_create(int n, const double *H, const double *T, const double *D, const double *M,
                     const double *K, const double *CP, void *stream, chs_props **out);
int chs_props_destroy(chs_props *p);

/* THERM_PROP_UPDATE (src/THERM_PROP_UPDATE.f90:129-205) on device arrays:
 * per cell T = S_HT(h), rho = S_TD(T), mu = S_TM(T), kappa = S_TK(T) and
 * c_p = S_TCp(T) (cp_dev may be NULL).  Cells with |h| < 1e-10 are left
 * unchanged, as the reference does.  The cells are nrow rows of rowlen
 * consecutive values, row r starting at r * row_stride (for ENTHALPY(NCL1,
 * 0:N2DO+1, NCL3): pass the address of (1,1,1), nrow = NCL3, rowlen =
 * NCL1*N2DO, row_stride = NCL1*(N2DO+2); for a flat array nrow = 1).
 * mode CHS_PROPS_TABLE: spline lookup.  mode CHS_PROPS_SURROGATE: the
 * trained MLP where it is certified, the table elsewhere; *n_fallback (may
 * be NULL) receives the number of cells the surrogate handed to the table.
 * Cells whose h or T leaves the table are left unchanged and the call
 * returns CHS_E_RANGE (the reference STOPs, src/Flow_State.f90:1030-1046).
 * Synchronous: returns after the update has finished. */
int chs_therm_prop_update(chs_props *p, int mode, const double *h_dev, long long nrow, long long rowlen,
                          long long row_stride, double *T_dev, double *rho_dev, double *mu_dev,
                          double *kap_dev, double *cp_dev, long long *n_fallback);

/* The form of the property update inside chs_energy_substep: CHS_PROPS_TABLE
 * (default, the reference's spline lookups) or CHS_PROPS_SURROGATE (the
 * trained, certified network for the final THERM_PROP_UPDATE of the cascade,
 * src/SOLVERRK3_ENG_IO.f90:46-66; the table for cells outside certified bins
 * and for ENTHALPY_UPDATE / DENSITY_UPDATE).  No counterpart in the
 * reference.  CHS_E_ARG if the surrogate is not trained and certified. */
int chs_props_set_mode(chs_props *p, int mode);
/* Cells of the last chs_energy_substep that the surrogate handed to the table.
 * Waits for that call's counts; returns its range error if not yet reported. */
int chs_props_last_fallback(chs_props *p, long long *n_fallback);

/* The range check of the last chs_energy_substep with p: waits for that
 * call's counts and returns CHS_E_RANGE if any property lookup left the table
 * (the reference STOPs in THERM_PROP_UPDATE, src/Flow_State.f90:1030-1046).
 * A range error is reported once, by the first of: this call,
 * chs_props_last_fallback, or the next chs_energy_substep with p.  No
 * counterpart in the reference: the energy substep reads its counts without a
 * stream sync, so that the host can queue the momentum substep meanwhile. */
int chs_props_check(chs_props *p);

/* RHOHmax of the table (NIST_SPLINE_H, src/Flow_State.f90:960-990): the
 * limiter of the energy update. */
int chs_props_rhoh_max(chs_props *p, double *rhoh_max);

/* ---- energy equation (variable-property heated channel) ------------------
 * Device arrays of the thermal state, reference layouts:
 *   (NCL1, 0:NCL2+1, NCL3): RHOH, ENTHALPY, TEMPERATURE, DENSITY, DENSITY0,
 *       DENSITY1, VISCOUSITY, VISCOUSITY0, THERMCONDT
 *   RHS_ENERGY0 (NCL1, NCL2, NCL3)
 *   G_io, G0_io (NCL1, 0:NCL2+1, NCL3, 3)   mass fluxes of this and the last substep
 *   WALLFLUX (NCL1, 2)                       wall heat flux, bottom then top */
typedef struct chs_thermal_fields {
    double *rhoh, *enthalpy, *temperature, *density, *density0, *density1;
    double *viscosity, *viscosity0, *thermcondt, *rhs_energy0;
    const double *g, *g0, *wallflux;
} chs_thermal_fields;

/* One energy substep SOLVERRK3_ENG_IO(NS) (src/SOLVERRK3_ENG_IO.f90:1-77) on
 * the grid of h (periodic x and z, iso-flux walls, table properties of p):
 * RHS_ENERGY_EXPLICIT (src/RHS_ENERGY_EXPLICIT.f90:1-150), the RK3 update of
 * RHOH with the RHOHmax limiter, ENTHALPY_UPDATE / DENSITY_UPDATE /
 * THERM_PROP_UPDATE (src/THERM_PROP_UPDATE.f90:1-205) and BC_WALL_THERMAL
 * (src/BC_WALL.f90:522-650).  cthecd = 1/(REN PRT0).  Work is ordered on h's
 * stream and asynchronous: the count of lookups that left the table is read
 * by the next call with p, which then returns CHS_E_RANGE before doing any
 * work (the reference STOPs), or by chs_props_check / chs_props_last_fallback.
 * Calls that share p are ordered by that wait. */
int chs_energy_substep(chs_handle *h, chs_props *p, const chs_thermal_fields *f, int ns, double dt,
                       double cthecd);

/* ---- io momentum substep of the heated channel ----------------------------
 * The main-domain (io) momentum equation with variable properties in mass-flux
 * form G = rho u, periodic x and z, channel walls, explicit viscous terms
 * (visthemflg = 0) and no weighted pressure.  Arrays in the reference's
 * layouts (x fastest):
 *   q, g, g0        Q_io, G_io, G0_io (NCL1, 0:NCL2+1, NCL3, 3)
 *   pr              PR_io (NCL1, 0:NCL2+1, NCL3)
 *   drhodtp, divu   DrhoDtP, DivU_io (NCL1, 0:NCL2+1, NCL3)
 *   rho_stg, mu_stg RHO_STG, MU_STG (NCL1, 0:NCL2+1, NCL3, 3)
 *   explt0          EXPLT0_io (NCL1, NCL2, NCL3, 3)
 * g and g0 are the arrays the energy substep reads (chs_thermal_fields.g/g0).
 * The properties come from chs_thermal_fields: density, density0,
 * viscosity, viscosity0 (used as given, ghost rows included). */
typedef struct chs_io_fields {
    double *q, *g, *g0, *pr, *drhodtp, *divu, *rho_stg, *mu_stg, *explt0;
    double *drhoi_stg;   /* DRHOI_STG (NCL1, 0:NCL2+1, NCL3, 3): visimplicit = 1 only (may be NULL otherwise) */
} chs_io_fields;

/* thermal: thermlflg (1 = variable properties: DrhoDtP, RHO_STG and MU_STG
 *          are computed; 0 = they are left as the caller set them, 0 and 1);
 * nflow: NFLOW 1..3; flowdrvtp: 1 constant mass flux, 2 constant pressure
 * gradient with cfgv = CFGV; gravdir: GRAVDIR 0..3 with f_a = F_A
 * (src/Flow_State.f90:97-110); visimplicit: visthemflg (src/READINI.f90:225). */
typedef struct chs_io_params {
    int thermal, nflow, flowdrvtp, gravdir;
    double cfgv, f_a;
    int visimplicit;     /* visthemflg: 0 explicit viscous terms, 1 RK-CN (VISCOUS_PAR_EXPLT_*_io + MOMFA_io) */
} chs_io_params;

/* What SOLVE runs between SOLVERRK3_ENG_IO(NS) and SOLVERRK3_MOM_IO(NS)
 * (src/SOLVE.f90:121-133): DENSITY_TIME_DERIVATION(NS)
 * (src/THERM_PROP_UPDATE.f90:225-291), DENSITY_Staggered, MU_Staggered and
 * VELOCITY_CALC_io (src/MASSFLUX_CALC_IO.f90:82-311).  ns = 1..3. */
int chs_io_prepare(chs_handle *h, const chs_thermal_fields *tf, const chs_io_fields *f, const chs_io_params *p,
                   int ns);
/* SOLVERRK3_MOM_IO(NS) (src/SOLVERRK3_MOM_IO.f90:1-168).  Explicit viscous
 * branch: CONVECTION_*_io, DIVG_U_io, VISCOUS_ALL_EXPLT_*_io,
 * RHS_MOM_EXPLICIT_io (src/RHS_MOM_EXPLICIT_io.f90:1-297), MASSFLUX_CALC_IO;
 * RK-CN branch (visimplicit = 1): VISCOUS_PAR_EXPLT_*_io, RHS_MOM_EXPLICIT_io
 * and MOMFA_io (src/MOMFA_io.f90:17-377) per component; then in both
 * DIVG_io (src/DIVG.f90:60-150), the FFT99 Poisson solve (the reference's
 * alternative for the io domain, src/SOLVERRK3_MOM_IO.f90:123), PRCALC_io and
 * MASSFLUX_UPDATE_IO (src/VELO_CORRECTION.f90:109-185).  Updates g, g0, pr,
 * divu, explt0.  Asynchronous on h's stream. */
int chs_solverrk3_mom_io(chs_handle *h, const chs_thermal_fields *tf, const chs_io_fields *f,
                         const chs_io_params *p, int ns, double dt);
/* DIVGCK_io (src/DIVGCK.f90:61-124): max |div G + DrhoDtP| and max |div Q|
 * over the interior (NaN propagates).  Either output may be NULL. */
int chs_divgck_io(chs_handle *h, const chs_io_fields *f, double *maxdiv_g, double *maxdiv_q);

/* xz-plane moments of the heated channel, PP_MEAN_ZX_FLOW_Xperiodic_IO
 * (src/processing/PP_MEAN_ZX.f90:299-800), from Q_io, G_io, PR_io and (thermal
 * = 1) TEMPERATURE, DENSITY, ENTHALPY, VISCOUSITY.  With the cell-centred
 * u_m = (Q(i)+Q(i+1))/2 (and likewise g_m = (G(i)+G(i+1))/2, v and w along j
 * and k), out[37 j + m] for row j = 0..NCL2-1 holds
 *   m = 0..3    <u>, <v>, <w>, <p>          U1xzL_io(J,1:4)
 *   m = 4..6    <g_m>                       G1xzL_io
 *   m = 7..9    <u_m p>                     UPxzL_io
 *   m = 10..15  <u_m u_n>, (m,n) = 11 12 13 22 23 33   U2xzL_IO
 *   m = 16..21  <u_m g_n>, same order       UGxzL_IO
 *   m = 22..30  <T> <rho> <h> <mu> <T^2> <rho^2> <h^2> <h rho> <h p>
 *               T1, D1, H1, M1, T2, D2, H2, DH, PHxzL_io
 *   m = 31..33  <h u_m>                     UHxzL_io
 *   m = 34..36  <h g_m>                     GHxzL_io
 * (columns 22..36 are zero with thermal = 0).  out_host (37*NCL2 doubles)
 * may be NULL: the moments then only stay on the device for chs_mean_t.
 * One block per row sums in a fixed order: deterministic. */
int chs_mean_xz_io(chs_handle *h, const chs_thermal_fields *tf, const chs_io_fields *f, int thermal,
                   double *out_host);

/* ---- learned property surrogate ------------------------------------------
 * The reference has no surrogate: these entries train and certify the network
 * that chs_therm_prop_update(..., CHS_PROPS_SURROGATE, ...) runs in place of
 * the table where certified.  Network: h -> (T, rho, mu, kappa, c_p),
 * 1 -> 64 -> 64 -> 5, tanh; weights held as pairs of BF16 values (hi + lo),
 * activations split the same way, BF16 matrix-core products with FP32
 * accumulation; parameters are a flat float vector of 4613:
 * w1[64] b1[64] W2[64][64] b2[64] W3[5][64] b3[5].
 *
 * Train with Adam (cosine learning-rate decay from lr to lr/100) on nsamp
 * device enthalpy samples labelled by the table; batch is a multiple of 64.
 * The trained range is [min h, max h] of the samples.  A new network must be
 * certified before use.  *final_loss: mean squared relative error over all
 * samples. */
int chs_surrogate_train(chs_props *p, const double *h_dev, long long nsamp, int steps, int batch, double lr,
                        unsigned long long seed, double *final_loss);
/* Split the trained range into nbins bins; a bin is certified when the
 * network is within tol/2 (relative, every output) of the table on 33 points
 * spanning it.  Uncertified bins and cells outside the range use the table. */
int chs_surrogate_certify(chs_props *p, double tol, int nbins, double *certified_fraction);
/* Fractions of the bins the last certification passed for all five outputs
 * and for T, rho, mu, kappa alone: a call that passes cp_dev = NULL (and the
 * energy substep in surrogate mode) uses the network in the bins of the
 * second kind. */
int chs_surrogate_certified(chs_props *p, double *all_outputs, double *without_cp);
/* Loss and gradient (host, 4613 floats) over exactly the n device samples
 * (n <= 8192), with no update: the training kernels' gradient, for checking. */
int chs_surrogate_grad(chs_props *p, const double *h_dev, long long n, float *grad_host, double *loss);
/* Raw network outputs, no fallback: out_dev[k*n + q], k = T, rho, mu, kappa, c_p. */
int chs_surrogate_eval(chs_props *p, const double *h_dev, long long n, double *out_dev);
/* op 0: read the parameters (and norm[14] = hc, hs, h_lo, h_hi, m[5], s[5]);
 * op 1: set them (count = 4613); x = (h - hc) hs, y_k = m_k + s_k o_k. */
int chs_surrogate_params(chs_props *p, int op, float *host, int count, double *norm);

/* Per-kernel timing with HIP events on the handle's stream. */
int chs_profile_enable(chs_handle *h, int on);
/* Fills up to cap entries; returns the number of kernels recorded in *count.
 * names: cap * 48 chars (NUL-terminated each); ms: total ms; calls: launches. */
int chs_profile_read(chs_handle *h, int cap, char *names, double *ms, long *calls, int *count);
int chs_profile_reset(chs_handle *h);

const char *chs_last_error(void);
const char *chs_version(void);

#ifdef __cplusplus
}
#endif

#endif /* CHAPSIM_GPU_H */
