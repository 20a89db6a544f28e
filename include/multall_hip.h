/* multall_hip.h -- C ABI of the MI355X time-marching library.
 *
 * Drop-in replacement for the reference solver's main iteration loop
 * (Multall-open-18.3/multall-18.3/multall-open-18.3.f):
 *
 *   - mh_start + mh_step replace `CALL LOOP` (mf:137, body mf:3664-5463):
 *     the prologue at mf:3674-3714 and the time steps from label 5000, each
 *     step doing the reference's primitive recovery, face fluxes, the five
 *     `TSTEP(D,DIFF,NCALL)` updates (mf:4251-5105, TSTEP mf:5470-5932 with
 *     SMOOTH_VAR mf:15352-15594 and SMOOTH_RESID mf:14833-14917), the
 *     inlet/exit boundary conditions (mf:4476-4757), the limiters
 *     (mf:5115-5168), NEW_MIXPLAN (mf:14922-15251), the every-5-step
 *     convergence checks (mf:5252-5317) and STEPUP(0.25) (mf:7916-8002);
 *   - mh_set_array / mh_get_array stand for the COMMON blocks the loop
 *     reads and writes (commall-open-18.3:13-203);
 *   - mh_integrals replaces SUMFLX (mf:7873-7912) as called by EFICOOL
 *     (mf:10741), accumulating in FP64;
 *   - mh_eficool replaces EFICOOL's stage and machine performance
 *     (mf:11253-11548), in FP64.
 *
 * Conventions: arrays use the reference's names and its I-fastest,
 * column-major order.  Dense arrays are packed to the logical extents
 * (IM, JM, KM); mh_set_array_strided accepts the padded COMMON arrays
 * (leading dimensions ID, JD).  Integer index maps keep the reference's
 * 1-based values.  Every entry point returns MH_OK (0) or a negative error
 * code; mh_last_error() gives the message.  A context is not thread-safe;
 * it owns its device buffers and HIP stream.  No host pointer is retained
 * after a call returns.
 */
#ifndef MULTALL_HIP_H
#define MULTALL_HIP_H
#ifdef __cplusplus
extern "C" {
#endif

#define MH_OK 0
#define MH_E_ARG -1      /* bad argument, unknown id, size mismatch */
#define MH_E_HIP -2      /* a HIP runtime call failed */
#define MH_E_UNSUPPORTED -3 /* an option the library does not implement */
#define MH_E_STATE -4    /* call out of order (e.g. mh_step before mh_start) */
#define MH_E_NAN -5      /* NaN/Inf detected in the convergence reductions */

typedef struct mh_ctx mh_ctx;

/* Problem extents (reference names).  nib1.. are the multigrid block counts
 * NIB1, NJB1, NKB1, NIB2, NJB2, NKB2 and the superblock count NSBLK
 * (commall-open-18.3:89-93), as set by the reference's SETUP (mf:7067-7180). */
typedef struct mh_dims {
  int im, jm, km, nrows;
  int nib1, njb1, nkb1, nib2, njb2, nkb2, nsblk;
} mh_dims;

/* Convergence summary of the last every-5-step check (mf:5252-5317). */
typedef struct mh_resid {
  int nstep;        /* steps taken since mh_start (NSTEP) */
  int ifend;        /* 1 once the run has ended (NSTEPS_MAX, converged) */
  float emax, eavg, econt, vref, vmax;
  float flow_in, flow_out; /* BLADE_FLOW(1), BLADE_FLOW(JM), kg/s */
  int imax, jmax, kmax;     /* where the last check found EMAX (IMAX, JMAX, KMAX; 0 before any) */
} mh_resid;

/* Array ids.  Shapes: N = (IM,JM,KM) node/cell arrays; JK = (JM,KM);
 * IK = (IM,KM); I, K, J = vectors; R = per row (NROWS+1); RN = per row
 * (NROWS); B1/B2/SB =
 * multigrid step arrays STEP1 (NIB1,NJB1+1,NKB1), STEP2, STEPSBK. */
enum mh_array_id {
  /* conserved variables and SSS history (BKPRIM, BKDIFF) -- N */
  MH_RO = 0, MH_ROVX, MH_ROVR, MH_RORVT, MH_ROE,
  MH_DRO, MH_DROVX, MH_DROVR, MH_DRORVT, MH_DROE,
  /* secondary variables (BKSEC, BKPRIM) -- N */
  MH_ROVT, MH_ROWT, MH_P, MH_HO, MH_PEFF, MH_VX, MH_VR, MH_VT, MH_WT, MH_T_STATIC,
  /* face areas, volumes, time steps (BKAREA, BKFLUX) -- N */
  MH_AQX, MH_AQR, MH_ABX, MH_ABR, MH_ASX, MH_ASR, MH_VOL, MH_VOLOR,
  MH_STEP, MH_BSTEP, MH_RSTEP,
  /* body forces (BKLOSS) -- N */
  MH_XFORCE, MH_RFORCE, MH_TFORCE, MH_QSOURCE,
  /* face mass fluxes (BKFLOW) -- N */
  MH_FLOWX, MH_FLOWT, MH_FLOWR,
  /* (J,K) geometry -- JK */
  MH_R, MH_X, MH_RAVG_CELL, MH_ABT, MH_WRABT, MH_UBLADE, MH_DX, MH_DR, MH_DS, MH_SMERID,
  /* corner distribution weights (BKDIST) -- IK */
  MH_FTL, MH_FBL, MH_FTR, MH_FBR,
  /* pitchwise vectors -- I */
  MH_FP, MH_FU, MH_FD,
  /* spanwise vectors -- K */
  MH_FR, MH_FKU, MH_FKD, MH_PO1, MH_TO1, MH_RPO1, MH_BSSIN, MH_BSCOS, MH_BRSIN, MH_BRCOS,
  MH_VTIN, MH_PD, MH_VM1,
  /* streamwise vectors -- J */
  MH_FACUP, MH_FACDWN, MH_WRAD,
  /* multigrid block time steps -- B1, B2, SB */
  MH_STEP1, MH_STEP2, MH_STEPSBK,
  /* integer maps (1-based values) */
  MH_IB1, MH_IB2,                                   /* I */
  MH_KB1, MH_KB2,                                   /* K */
  MH_NBLADE, MH_IND, MH_INDLE, MH_INDMIX, MH_INDTE, MH_NROW, MH_JSBLK, MH_JB1, MH_JB2, /* J */
  MH_JSTART, MH_JMIX, MH_JLE, MH_JTE, MH_JLED, MH_KTIPS, MH_KTIPE, MH_KS1, MH_KS2,  /* R */
  /* mixing-length viscous model, LOSS (BKLOSS, BKNEWLOS, BKTIP): squared
   * cell wall distance DWALLSQ (SET_XLENGTH, mf:12498-12508) and the y+ it
   * yields -- N; end-wall rotation rates -- J; per-row mixing-length limits,
   * wall roughnesses (m) and transition points -- RN = NROWS entries */
  MH_DWALLSQ, MH_Y_PLUS,                            /* N */
  MH_WHUB, MH_WTIP,                                 /* J */
  MH_XLLIM_I1, MH_XLLIM_IM, MH_XLLIM_K1, MH_XLLIM_KM, MH_XLLIM_DWN, MH_XLLIM_UP,
  MH_ROUGH_H, MH_ROUGH_T, MH_ROUGH_L, MH_ROUGH_U,   /* RN, float */
  MH_JTRAN_I1, MH_JTRAN_IM, MH_JTRAN_K1, MH_JTRAN_KM, /* RN, int */
  /* Spalart-Allmaras model, SPAL_LOSS (BKSPALOS, BKNEWLOS; NEW_LOSS uses
   * the cell density, laminar viscosity, viscosity ratio, mixing length and
   * free-stream factor): the transported
   * turbulent viscosity and its SSS history, its kinematic value and source,
   * the cell density, laminar viscosity and viscosity ratio, the mixing
   * length and free-stream factor of SET_XLENGTH -- N (cell-centred); the
   * free-stream turbulence ratio and mixing-plane decay per row -- RN */
  MH_TRANS_DYN_VIS, MH_DEL_DYNVIS, MH_TRANS_KVIS, MH_T_SOURCE, MH_ROAVG_CELL, MH_VISC_LAM, MH_VISC_RAT,
  MH_XLENGTH, MH_FST_RAT,
  MH_FSTURB, MH_TURBVIS_DAMP,
  /* Shih et al. wall functions (YPLUSWALL < -0.001; BKLOSS, BKNEWLOS): the
   * smoothed streamwise wall pressure gradient -- N (cell-centred); the y+
   * each wall cell's function iterates on -- JK for the I = 1 and I = IM
   * blade surfaces, IJ (I fastest, IM x JM) for the hub and casing.  They
   * start at SETUP's values (0 and 25, mf:6904-6910) when not set.  The
   * blade chord of each row, for the trailing-edge Reynolds number that
   * scales the SA pressure-gradient source factor -- RN */
  MH_DPDS_CELL, MH_YPLUS_I1, MH_YPLUS_IM, MH_YPLUS_K1, MH_YPLUS_KM, MH_CHORD,
  /* the five printout steps NOUT (BKCNTL, read at mf:723): a step that is
   * one of them and ends on a convergence check refreshes the viscous
   * forces there, as the reference does before OUTPUT (mf:3767-3769,
   * 5338-5341) -- 5 ints; unset (0) means none inside the run */
  MH_NOUT,
  /* the stage number of each blade row (BKSTAG NSTAGE, read at mf:710-711)
   * -- RN ints; unset means the reference's default 1 + (N-1)/2 (mf:704-706).
   * Used only by mh_eficool. */
  MH_NSTAGE,
  /* the trailing-edge separation force of rows with IF_CUSP = 2 (BKSTAG and
   * the reader's per-row inputs, mf:993-1004; applied each step,
   * mf:5184-5233): the per-row option, the points upstream of the trailing
   * edge where it starts on the I = 1 and I = IM surfaces, the wake points
   * it extends over -- RN ints; the layer thickness as a fraction of the
   * blade thickness and the momentum factor -- RN floats; the blade
   * tangential thickness RT_THICK -- JK; the node pitchwise positions
   * R*THETA (RTHETA) -- N.  Needed only when some row has IF_CUSP = 2. */
  MH_IF_CUSP, MH_NUP_I1, MH_NUP_IM, MH_N_WAKE, MH_SEP_THIK, MH_SEP_DRAG, MH_RT_THICK, MH_RTHETA,
  /* the inlet yaw and pitch angles in degrees (BS, BR; BKINPT) -- K; the
   * repeating-stage condition relaxes BS (mf:12100) and recomputes the
   * sines and cosines from both.  Needed only with IF_REPEAT != 0. */
  MH_BS, MH_BR,
  MH_NARRAYS
};

/* Scalar ids (reference names; BKSMTH, BKCNTL, BKGAS, BKINPT, BKSTAG ...). */
enum mh_scalar_id {
  MH_S_CFL = 0, MH_S_DAMPIN, MH_S_MACHLIM, MH_S_F_PDOWN, MH_S_CONLIM, MH_S_SFXIN, MH_S_SFTIN,
  MH_S_FAC_4TH, MH_S_F1, MH_S_F2, MH_S_F3, MH_S_RSMTH, MH_S_SFEXIT, MH_S_FSMTHB, MH_S_RFIN,
  MH_S_PDOWN_HUB, MH_S_PDOWN_TIP, MH_S_FEXTRAP, MH_S_RFMIX, MH_S_FANGLE, MH_S_CP, MH_S_GA,
  MH_S_RGAS, MH_S_CV, MH_S_FGA, MH_S_GA1, MH_S_RCP, MH_S_VSOUND, MH_S_FACMIXUP,
  MH_S_IMID, MH_S_KMID, MH_S_NSTEPS_MAX, MH_S_NCHANGE, MH_S_NRSMTH, MH_S_NSFEXIT, MH_S_IN_PRESS,
  MH_S_IN_VTAN, MH_S_IN_VR, MH_S_IPOUT, MH_S_IFMIX, MH_S_IR, MH_S_JR, MH_S_KR, MH_S_ILOS,
  MH_S_NLOS, MH_S_NMIXUP, MH_S_ITIMST, MH_S_IF_RESTART, MH_S_IFGAS, MH_S_IN_FLOW,
  MH_S_IF_REPEAT, MH_S_IRBB, MH_S_KRBB,
  /* viscous model (BKLOSS, BKGAS, BKNEWLOS) */
  MH_S_REYNO, MH_S_RF_VIS, MH_S_FTRANS, MH_S_PRANDTL, MH_S_YPLUSWALL, MH_S_YPLAM, MH_S_YPTURB,
  MH_S_CFWALL, MH_S_IBOUND,
  /* Spalart-Allmaras source factors and limits (BKSPALOS) */
  MH_S_FAC_STMIX, MH_S_FAC_ST0, MH_S_FAC_ST1, MH_S_FAC_ST2, MH_S_FAC_ST3, MH_S_FAC_SFVIS, MH_S_FAC_VORT,
  MH_S_FAC_PGRAD, MH_S_VORT_MAX, MH_S_TURBVIS_LIM,
  /* mass-flow forcing of SETFLO (BKINPT): target inlet flow and relaxation
   * factor, read only when IN_FLOW != 0 (required then, default 0) */
  MH_S_FLOWIN, MH_S_RFLOW,
  /* the repeating-stage condition (IF_REPEAT != 0; required then, default
   * 0): steps between updates and relaxation factor (mf:575-585), and the
   * deck's mid-span inlet stagnation pressure and tangential velocity
   * (PO1 and VTIN at the middle input point, mf:1889, 1911) */
  MH_S_NINMOD, MH_S_RFINBC, MH_S_PO_IN_MID, MH_S_VT_IN_MID,
  /* optional state: laminar viscosity and conductivity ratio left by the
   * last LOSS call (default 0; mh_loss_init sets them) */
  MH_S_VISLAM, MH_S_FTCOND,
  /* library controls (no reference counterpart) */
  MH_S_EXACT_ORDER,   /* 1: reductions in the reference's sequential order */
  MH_S_USE_GRAPHS,    /* 1 (default): steps queued as captured HIP graphs, 0: launch by launch */
  MH_S_LANES,         /* 1 (default): in the fast mode, the time steps of independent equations batched in one
                         launch per kernel (continuity, energy, turbulence; the three momentum equations) */
  MH_S_MIXED,         /* 1: the mixed form (fast mode only, read at mh_start): the SSS histories DRO..DROE and
                         DEL_DYNVIS stored in BF16, all arithmetic in FP32; 0 (default): FP32 storage */
  MH_S_FLOW_CONV,     /* > 0: also end the run once the inlet mass flow has changed by less than this
                         fraction between each two of the last 10 every-5-step checks (50 steps) and
                         ECONT < 0.01 (default 0: off; the reference's own stop tests always apply) */
  MH_S_TUNE,          /* 1 (default): at mh_start, time the tile shapes of the damping/corner-distribution
                         and SMOOTH_VAR kernels on this grid and keep the fastest (remembered per grid
                         shape for the process); 0: the fixed shapes; every shape gives the same bits */
  MH_S_SCAT_TK,       /* 2, 4 or 8: force the damping/corner tile's K lines (0, default: the tuner's) */
  MH_S_SMOOTH_WAVES,  /* 8 or 16: force SMOOTH_VAR's wavefronts per J plane (0, default: the tuner's) */
  MH_NSCALARS
};

/* Create a context for one machine (all rows of one deck). */
int mh_create(const mh_dims *dims, mh_ctx **out);
int mh_destroy(mh_ctx *ctx);
const char *mh_last_error(const mh_ctx *ctx);

/* Name lookups ("RO", "CFL" ...); return -1 if unknown. */
int mh_array_id(const char *name);
int mh_scalar_id(const char *name);
/* Number of elements of an array for this context's extents. */
long long mh_array_len(const mh_ctx *ctx, int id);
/* 1 for int32 arrays, 0 for float arrays, -1 for a bad id. */
int mh_array_is_int(int id);

/* Copy a dense host array (float or int32 per mh_array_is_int) in/out. */
int mh_set_array(mh_ctx *ctx, int id, const void *host, long long n);
int mh_get_array(mh_ctx *ctx, int id, void *host, long long n);
/* Same, from/to a padded COMMON array with leading dimensions ld_i (ID) and
 * ld_j (JD) for N arrays, ld_j (JD) for JK arrays, ld_i (ID) for IK and IJ
 * arrays. */
int mh_set_array_strided(mh_ctx *ctx, int id, const void *host, int ld_i, int ld_j);
int mh_get_array_strided(mh_ctx *ctx, int id, void *host, int ld_i, int ld_j);
int mh_set_scalar(mh_ctx *ctx, int id, double value);
double mh_get_scalar(const mh_ctx *ctx, int id);

/* The viscous-force initialisation SETUP performs before LOOP
 * (mf:7346-7357): one LOSS call (mf:8058-9066) with RF_VIS = 1, FMIXUP = 1
 * and NSTEP = 1, which also sets VISLAM/FTCOND; with 100 <= ILOS < 200 the
 * call is NEW_LOSS (mf:12522-13462), with ILOS >= 200 SPAL_LOSS
 * (mf:13586-14827), which also sets the initial turbulent viscosity.  A
 * no-op for ILOS = 0.  Call after all inputs are set, before
 * or after mh_start. */
int mh_loss_init(mh_ctx *ctx);

/* SET_XLENGTH (multall-open-18.3.f:12194-12517) on the device: the wall
 * distances and mixing-length scales DWALLSQ, XLENGTH and FST_RAT from the
 * grid already set (X, R, SMERID, ASX, ASR, ABX, ABR, ABT, the row tables,
 * FSTURB and every scalar), instead of passing them in.  theta: the nodal
 * THETA(I,J,K) of the reference's COMMON, dense (IM,JM,KM), I fastest;
 * xllim_in/le/te/dn: the per-row mixing-length limits XLLIM_IN, XLLIM_LE,
 * XLLIM_TE, XLLIM_DN, NROWS
 * values each (XLLIM_* read at mf:2072).  The reference calls SET_XLENGTH from
 * SETUP (mf:7340); the replacement is one call after the uploads and before
 * mh_loss_init.  The values equal the reference's within float rounding (the
 * device's sinf and powf). */
int mh_set_xlength(mh_ctx *ctx, const float *theta, const float *xllim_in, const float *xllim_le,
                   const float *xllim_te, const float *xllim_dn);

/* mh_set_xlength without allocating: the same computation (the same
 * kernels, bit-identical results), its work buffers in a file-scope device
 * array of the library instead of device memory allocated per call.  No
 * counterpart in the reference; for callers that cannot allocate device
 * memory (under HIP graph capture, in a serving loop, in harnesses that
 * forbid hipMalloc).  It calls no hipMalloc/hipFree.  Limits: at most
 * MH_XLENGTH_SCRATCH_NODES nodes (IM*JM*KM) and MH_XLENGTH_SCRATCH_J
 * streamwise points; larger grids get MH_E_ARG (use mh_set_xlength).  The
 * array (about 32 MiB) is held on each device the library runs on until the
 * process ends and is shared by all contexts on that device: calls that
 * overlap on one GPU race, so the caller serialises them. */
#define MH_XLENGTH_SCRATCH_NODES 4194304
#define MH_XLENGTH_SCRATCH_J 65536
int mh_set_xlength_noalloc(mh_ctx *ctx, const float *theta, const float *xllim_in, const float *xllim_le,
                           const float *xllim_te, const float *xllim_dn);

/* LOOP prologue (mf:3674-3714).  Call once after all inputs are set. */
int mh_start(mh_ctx *ctx);
/* Run up to nsteps time steps (fewer if the run ends).  *out (optional)
 * receives the latest convergence summary. */
int mh_step(mh_ctx *ctx, int nsteps, mh_resid *out);
/* A stop request, as the reference's `stopit` file (read at the start of
 * every step N with N divisible by 10, mf:3774-3779): the next such step
 * becomes the run's last (IFEND = 1), with its closing viscous call and
 * convergence check, and NSTEPS_MAX keeps its value.  The host polls the
 * file between mh_step calls and calls this when it holds a value >= 1. */
int mh_request_stop(mh_ctx *ctx);

/* Several independent machines (operating points of a speedline, say) at
 * once: each of the n distinct contexts advances up to nsteps steps, as
 * mh_step would, driven by a host thread of its own on its own stream, so
 * the kernels of different contexts overlap on the GPU.  out[i] receives
 * context i's residuals.  Returns 0, or the first failing context's code
 * with its index in *failed (the message is that context's
 * mh_last_error).  No counterpart in the reference, which runs one
 * operating point per process; the drivers of a speedline or of
 * operating-point sweeps call it. */
int mh_step_batch(mh_ctx *const *ctxs, int n, int nsteps, mh_resid *out, int *failed);
/* Per-J mass-flux-weighted sums of SUMFLX (mf:7883-7910) in FP64:
 * out[7*JM] = {flow, sum_po, sum_to, sum_rvt, sum_entropy, sum_tstat,
 * sum_pstat} per J, with p0/po1 and T0 from the current state. */
int mh_integrals(mh_ctx *ctx, double *out, long long n);

/* Stage and machine performance from mass averages: EFICOOL's uncooled
 * branch (mf:11253-11548; IFCOOL = IFBLEED = 0, IFGAS = 0), in FP64 on the
 * device from the SUMFLX integrals of the current state.  One record per
 * stage (the rows of one NSTAGE value; mf:3600-3614), then one for the whole
 * machine (stage = 0) when there is more than one stage, as the reference
 * prints them.  A stage's planes are J1 = its first J and J2M1 = the J after
 * its last row's JMIX (JM - 1 for the last stage); the whole machine's are
 * J = 2 and JM - 1.  Quantities at a plane are mass averages (SUMFLX sums
 * over BLADE_FLOW); pressures in Pa, temperatures in K, flows in kg/s for
 * the whole annulus. */
typedef struct {
  int stage;            /* 1 .. NSTAGES; 0 = the whole machine */
  int j1, j2m1, jmix;   /* inlet, exit and mid-stage planes (1-based J) */
  int compressor;       /* 1 when the stagnation temperature rises (DELTO >= 0) */
  int pad_;
  double po_in, po_out;       /* stagnation pressure (SUMPO x PO1(KMID)) */
  double ps_in, ps_out;       /* static pressure */
  double to_in, to_out;       /* stagnation temperature */
  double flow_in, flow_out;   /* mass flow at J1 and J2M1 (the machine: J = 1 and JM - 1) */
  double po_ratio;            /* po_out / po_in, not below 0.01 (PORAT) */
  double eta_tt, eta_ts;      /* total-to-total and total-to-static isentropic efficiency */
  double eta_poly;            /* total-to-total polytropic efficiency */
  double reaction;            /* from the mass-averaged static temperatures at J1, JMIX, J2M1 (stages) */
  double reaction_hub, reaction_mid, reaction_tip; /* from FP-weighted pitchwise static pressures at
                                                      K = 1, KM/2, KM (stages) */
  double work_kw;             /* CP (flow_in to_in - flow_out to_out) / 1000, kW */
} mh_perf;
/* Fills out[0 .. *n-1]; MH_E_ARG when cap < the number of records (then *n
 * still says how many there are). */
int mh_eficool(mh_ctx *ctx, mh_perf *out, int cap, int *n);

/* MIX_BCONDS (mf:11906-12080): the spanwise profiles of mass-averaged flow
 * at the inlet (J = 1) and at each row's mixing plane or exit (J = JMIX),
 * as the reference writes them to 'mixbconds' and the repeating-stage
 * condition uses them.  Per plane p (0 = inlet, p = row number) and
 * quantity q (0 stagnation pressure, 1 static pressure, 2 stagnation
 * temperature, 3 static temperature, 4 absolute tangential velocity,
 * 5 meridional velocity, 6 yaw angle, 7 pitch angle, both in degrees):
 * out[(p*8 + q)*(KM+1) + K-1] for K = 1..KM (the cell-line mass averages
 * moved to the nodes as MASS_AVG does, mf:11729-11760), and
 * out[(p*8 + q)*(KM+1) + KM] the whole plane's mass average.  Weights are
 * the cells' positive mass flows; sums in FP64.  n = (NROWS+1)*8*(KM+1). */
int mh_mix_bconds(mh_ctx *ctx, double *out, long long n);
/* Wait for all queued work; returns MH_E_HIP on an asynchronous error. */
int mh_sync(mh_ctx *ctx);

/* Kernel timing (no counterpart in the reference, whose only timing is the
 * MCLOCK CPU time per step, mf:5404-5414).  When enabled, every kernel launch
 * is bracketed by HIP events recorded on the context's stream.  The report
 * holds one line per kernel: "<name> <launches> <total milliseconds>". */
int mh_profile_enable(mh_ctx *ctx, int on);
int mh_profile_report(mh_ctx *ctx, char *buf, int cap);

#ifdef __cplusplus
}
#endif
#endif
