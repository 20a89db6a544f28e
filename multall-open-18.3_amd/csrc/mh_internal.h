// Internal definitions shared by the HIP translation units of libmultall_hip.
//
// Layout in HBM: every node/cell array is one dense FP32 buffer [KM][JM][IM]
// (I fastest, so a pitchwise line of IM floats is one coalesced segment);
// (J,K) arrays are [KM][JM]; (I,K) arrays are [KM][IM].  Cell (I,J,K) spans
// nodes I..I+1, J-1..J, K..K+1 exactly as in the reference, and cell arrays
// share the node arrays' shape (the unused I=IM, K=KM, J=1 slots are kept so
// that indices match the reference one-for-one).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "multall_hip.h"

namespace mh {

enum Kind : int { K_N, K_JK, K_IK, K_IJ, K_I, K_K, K_J, K_R, K_B1, K_B2, K_SB, K_RN, K_NOUT };

struct ArrayInfo {
  const char *name;
  Kind kind;
  bool is_int;
};

extern const ArrayInfo kArrays[MH_NARRAYS];
extern const char *const kScalarNames[MH_NSCALARS];

// Extents and derived counts passed by value to every kernel.
struct Geo {
  int IM, JM, KM, IMM1, JMM1, KMM1, IMID, KMID, NROWS;
  int NIB1, NJB1, NKB1, NIB2, NJB2, NKB2, NSBLK;
};

// Device pointers of every array, by reference name, plus scratch buffers.
struct Dev {
  float *RO, *ROVX, *ROVR, *RORVT, *ROE, *DRO, *DROVX, *DROVR, *DRORVT, *DROE;
  float *ROVT, *ROWT, *P, *HO, *PEFF, *VX, *VR, *VT, *WT, *T_STATIC;
  float *AQX, *AQR, *ABX, *ABR, *ASX, *ASR, *VOL, *VOLOR, *STEP, *BSTEP, *RSTEP;
  float *XFORCE, *RFORCE, *TFORCE, *QSOURCE, *FLOWX, *FLOWT, *FLOWR;
  float *R, *X, *RAVG_CELL, *ABT, *WRABT, *UBLADE, *DX, *DR, *DS, *SMERID;
  float *FTL, *FBL, *FTR, *FBR;
  float *FP, *FU, *FD;
  float *FR, *FKU, *FKD, *PO1, *TO1, *RPO1, *BSSIN, *BSCOS, *BRSIN, *BRCOS, *VTIN, *PD, *VM1;
  float *FACUP, *FACDWN, *WRAD;
  float *STEP1, *STEP2, *STEPSBK;
  int *IB1, *IB2, *KB1, *KB2, *NBLADE, *IND, *INDLE, *INDMIX, *INDTE, *NROW, *JSBLK, *JB1, *JB2;
  int *JSTART, *JMIX, *JLE, *JTE, *JLED, *KTIPS, *KTIPE, *KS1, *KS2;
  float *DWALLSQ, *Y_PLUS, *WHUB, *WTIP;
  float *XLLIM_I1, *XLLIM_IM, *XLLIM_K1, *XLLIM_KM, *XLLIM_DWN, *XLLIM_UP;
  float *ROUGH_H, *ROUGH_T, *ROUGH_L, *ROUGH_U;
  int *JTRAN_I1, *JTRAN_IM, *JTRAN_K1, *JTRAN_KM;
  float *TRANS_DYN_VIS, *DEL_DYNVIS, *TRANS_KVIS, *T_SOURCE, *ROAVG_CELL, *VISC_LAM, *VISC_RAT, *XLENGTH, *FST_RAT;
  float *FSTURB, *TURBVIS_DAMP;
  float *DPDS_CELL, *YPLUS_I1, *YPLUS_IM, *YPLUS_K1, *YPLUS_KM, *CHORD;
  int *NOUT;
  int *NSTAGE;
  int *IF_CUSP, *NUP_I1, *NUP_IM, *N_WAKE;
  float *SEP_THIK, *SEP_DRAG, *RT_THICK, *RTHETA;
  float *BS, *BR;
  double *MIXB;  // MIX_BCONDS profiles: (NROWS+1) planes x 8 quantities x (KM+1)
  // scratch (not part of the public array set)
  float *XFLUX, *TFLUX, *RFLUX, *SOURCE, *STORE, *DSC, *TMP, *TEMP1;
  float *B1CHG, *B2CHG, *SBCHG;      // multigrid block sums (FP32 results)
  float *BCW;                        // inlet/exit pitch averages per K, SA inlet density, swirl sums (4*KM+4)
  float *LTAB;                       // LOSS mixing-length factors, see LT_* below
  // Spalart-Allmaras work arrays (cell-centred, N): the velocity and
  // temperature gradients of SPAL_LOSS (strain rates, before the viscosity),
  // vorticity, turbulent and mixing-length viscosities, the gradient of the
  // kinematic turbulent viscosity, and the force/energy accumulators
  float *SXX, *SXR, *SXT, *SRR, *SRT, *STT, *GQX, *GQR, *GQT;
  float *ABS_VORT, *VISC_TURB, *DYN_VIS_ML, *GRAD_TVIS, *TVGX, *TVGR, *TVGT;
  float *FORCEX, *FORCER, *FORCET, *ESOURCE;
  int *SAFLAG;                       // transition flags per (I,J) column and (J,K) line
  int *JMP;                          // per J: mixing-plane role (low 2 bits) and JMIX of its interface
  int *SAJMAP;                       // per J: the J-1 and row the reference's periodic-force block reads
  double *RED;                       // FP64 partial sums of the reductions
  float *SC;                         // device-side scalars (see SC_* below)
  int *FLAGS;                        // per (I,K) line flags for the density limiter
  float *CHECK_FLOW, *BLADE_FLOW;    // per J
  float *SETFLO;                     // per J: SETFLO's FLOW(J), then its DELTA(J)
  float *AVG;                        // per row: the damping's mean |change| of the current TSTEP call
  float *KWALL;                      // kKWall planes of (IM-1) x (JM-1) end-wall column terms
  double *INTEG;                     // 7*JM FP64 integrals
  int *CTL;                          // device-resident step control (see CTL_* below)
  int gate;                          // CTL slot whose non-zero value makes a kernel return at once (-1: none)
};

// Slots of Dev::CTL, the step bookkeeping of LOOP (mf:3753-3824) kept on the
// device so that steps can be queued (and captured in graphs) without the
// host reading the convergence figures back.
enum {
  CTL_NSTEP = 0,   // NSTEP
  CTL_IFEND,       // IFEND of the current step
  CTL_DONE,        // the run ended at an earlier step (or on a non-finite check)
  CTL_SKIP,        // this step is not taken (the run has ended)
  CTL_NOTEND,      // SKIP, or this step is not the run's last: gates launches that only the last step makes
  CTL_STOPREQ,     // a stop request (stopit) is pending
  CTL_NANSTEP,     // the step whose convergence check found a non-finite value (0: none)
  CTL_FLOWOK,      // consecutive checks whose inlet-flow change met FLOW_CONV
  CTL_IMAX,        // where the last check found EMAX (1-based I, J, K)
  CTL_JMAX,
  CTL_KMAX,
  CTL_COUNT = 16
};

// A kernel launched as part of a time step returns at once when its gate is
// set (the run has ended, or a launch that only a run's last step makes).
__device__ __forceinline__ bool gated(const Dev &d) { return d.gate >= 0 && d.CTL[d.gate] != 0; }

// SA/NEW_LOSS end-wall terms per (I,J) column (forces, energy, friction
// velocity and density for hub and casing), and the cells of a column one
// k_sa_kfaces thread covers.
constexpr int kKWall = 12;
constexpr int kKSeg = 4;

// Slots of Dev::SC, the step's device-resident scalars.
enum {
  SC_VRMS = 0, SC_VMAX, SC_EMAX, SC_EAVG, SC_ECONT, SC_VREF, SC_FLOWRAT, SC_FLOWOUT, SC_VISLAM, SC_FTCOND,
  SC_VORTFAC, SC_PGRADFAC,  // SPAL_LOSS's helicity and pressure-gradient source factors
  SC_SFT, SC_SFX, SC_DAMP,  // the step's smoothing and damping factors (ramped over NCHANGE steps)
  SC_FMIXUP,                // the step's FMIXUP (squared), which the viscous calls read
  SC_FLOWPREV,              // inlet mass flow at the previous check
  SC_DFLOW,                 // its relative change between the last two checks (FLOW_CONV stop test)
  SC_NHEAD = 32,            // scalar slots before the per-row and per-K tables
  SC_AVG_CHG = SC_NHEAD,            // [NROWS]
  SC_AVG_BLK = SC_NHEAD + 64,       // [NROWS]
  SC_VLIMIT = SC_NHEAD + 128,       // [NROWS]
  SC_ROLIMIT = SC_NHEAD + 192,      // [NROWS]
  SC_PLIM = SC_NHEAD + 256,         // [KM], PLIM(K) fixed at the start of LOOP (mf:3684-3686)
  SC_COUNT = SC_NHEAD + 256         // + KM
};

// Layout of Dev::LTAB, the per-row mixing-length factors of LOSS
// (mf:8125-8183): FILAM, FITURB, FIWAKE, FIUP as [NROWS][IM] (pitch tables
// 0-3), then FKLAM, FKTURB as [NROWS][KM] (span tables 0-1).
enum { LT_FILAM = 0, LT_FITURB, LT_FIWAKE, LT_FIUP };
enum { LT_FKLAM = 0, LT_FKTURB };
__host__ __device__ inline int lt_pitch(int tab, int NR, int I, int NROWS, int IM) {
  return (tab * NROWS + NR - 1) * IM + I - 1;
}
__host__ __device__ inline int lt_span(int tab, int NR, int K, int NROWS, int IM, int KM) {
  return 4 * NROWS * IM + (tab * NROWS + NR - 1) * KM + K - 1;
}

// Scalars fixed for a run (set at the start of LOOP) handed to the kernels;
// those that change from step to step live on the device (SC_SFT ..).
struct StepScalars {
  float PREFF, ROLIM, TCOOL_MIN, SFEX1, SFEXH, RFIN1, RFIN2;
};

// Copy of the scalar table used by the kernels (floats and ints).
struct Params {
  float CFL, DAMPIN, MACHLIM, F_PDOWN, CONLIM, SFXIN, SFTIN, FAC_4TH, F1, F2, F3, RSMTH, SFEXIT,
      FSMTHB, RFIN, PDOWN_HUB, PDOWN_TIP, FEXTRAP, RFMIX, FANGLE, CP, GA, RGAS, CV, FGA, GA1, RCP,
      VSOUND, FACMIXUP;
  int IMID, KMID, NSTEPS_MAX, NCHANGE, NRSMTH, NSFEXIT, IN_PRESS, IN_VTAN, IN_VR, IPOUT, IFMIX, IR,
      JR, KR, ILOS, NLOS, NMIXUP, ITIMST, IF_RESTART, IFGAS, IN_FLOW, IF_REPEAT, IRBB, KRBB, EXACT;
  float REYNO, RF_VIS, FTRANS, PRANDTL, YPLUSWALL, YPLAM, YPTURB, CFWALL;
  int IBOUND;
  float FAC_STMIX, FAC_ST0, FAC_ST1, FAC_ST2, FAC_ST3, FAC_SFVIS, FAC_VORT, FAC_PGRAD, VORT_MAX, TURBVIS_LIM;
  float FLOWIN, RFLOW;
  int NINMOD;
  float RFINBC, PO_IN_MID, VT_IN_MID;
  float FLOW_CONV;  // library stop test on the inlet mass flow's relative change per check (0: off)
};

// Host-built J membership lists (CSR, 1-based J values) of the multigrid
// blocks (JB1, JB2 incl. the mixing-plane dummy block), the superblocks
// (JSBLK) and the row interiors used by the damping average.
// Node tile of the damping + corner-distribution kernel (TJ x TK node
// lines, all I; TK is tuned per grid from kScatTKs) and its dynamic LDS:
// cell lines (J..J+TJ, K-1..K+TK-1) and the four corner-weight arrays on
// K-1..K+TK-1, each `pitch` = 64 NC floats (NC = 1, 2 or 4 for IM-1 <= 64, 128, 256)
// plus a zero pad column at each end.
constexpr int kScatTJ = 4;
constexpr int kScatTKs[3] = {2, 4, 8};
inline int scatter_chunks(int imm1) { return imm1 <= 64 ? 1 : imm1 <= 128 ? 2 : 4; }
inline size_t scatter_lds_bytes(int imm1, int tk) {
  return (size_t)((kScatTJ + 1) * (tk + 1) + 4 * (tk + 1)) * (64 * scatter_chunks(imm1) + 2) * sizeof(float);
}
// SMOOTH_VAR's register form: wavefronts per workgroup (one J plane per
// workgroup), tuned per grid from kSmoothWaves.
constexpr int kSmoothWaves[2] = {8, 16};

// Fast-mode multigrid sums: per-(level-1 I block, J, level-1 K block) sums
// P and per-(J, level-1 K block) line totals PT, with each level-2 block's
// members as offsets into P.
struct MgFast {
  const int *b2_ptr = nullptr, *b2_off = nullptr;  // level-2 block -> offsets into P of its members
  double *P = nullptr, *PT = nullptr;
  const int *gate = nullptr;                       // the launch's gate flag (null: none)
};

// One conserved variable of a batched TSTEP launch: the variable, its SSS
// history (FP32, or its BF16 copy in the mixed form), its face fluxes and
// source, and its own TSTEP scratch.  A batched launch covers up to kMaxEq
// variables whose time steps do not read each other's results, one grid
// z-slice each.
struct EqLane {  // SOURCE may be null: no source term
  float *D, *DIFF;
  unsigned short *DIFF16, *STORE16;  // the mixed form's BF16 history and STORE
  float *XFLUX, *TFLUX, *RFLUX, *SOURCE, *STORE, *DSC, *TMP, *B1CHG, *B2CHG, *SBCHG, *AVG;
  double *RED, *P, *PT;
  int ncall;
};
constexpr int kMaxEq = 3;
struct EqSet {
  EqLane q[kMaxEq];
};

struct Lists {
  int *b1_ptr = nullptr, *b1_list = nullptr, *b2_ptr = nullptr, *b2_list = nullptr;
  int *sb_ptr = nullptr, *sb_list = nullptr, *row_ptr = nullptr, *row_list = nullptr;
  MgFast mgf;             // fast-mode multigrid sums (valid when mgf_ok)
  bool mgf_ok = false;
  float *mixscratch = nullptr;
};

}  // namespace mh

// One timed launch: HIP events recorded on the context's stream around it.
struct mh_prof_rec {
  const char *name;
  hipEvent_t t0, t1;
};

namespace mh {
// doubles per record of k_eficool (mh_eficool): po in/out, ps in/out, to
// in/out, flow in/out, PORAT, eta tt/ts/poly, reaction, reactions at hub,
// mid and tip, work, compressor flag
constexpr int kEffFields = 18;
constexpr int kHist = 6;  // history arrays with a BF16 copy in the mixed form
constexpr int kAhead = 8; // chunks (of up to 5 steps) the host may queue ahead of the device
// The flux and TSTEP scratch of a batched launch's second or third variable
// (time steps that do not read each other's results: continuity, energy and
// the turbulence transport; the three momentum equations).
struct Lane {
  unsigned short *STORE16 = nullptr;  // the mixed form's BF16 STORE
  float *XFLUX = nullptr, *TFLUX = nullptr, *RFLUX = nullptr, *SOURCE = nullptr, *STORE = nullptr, *DSC = nullptr,
        *TMP = nullptr, *B1CHG = nullptr, *B2CHG = nullptr, *SBCHG = nullptr, *AVG = nullptr;
  double *RED = nullptr;
  MgFast mgf;
};
constexpr int kLanes = 2;
}  // namespace mh

struct mh_ctx {
  bool prof_on = false;
  std::vector<mh_prof_rec> prof_pending;      // launches not yet reduced
  std::vector<hipEvent_t> prof_pool;          // recycled events
  std::vector<std::pair<std::string, std::pair<long long, double>>> prof_tot;  // name -> (calls, ms)
  mh::Lists lists{};
  size_t lds_plane = 0;
  mh_dims dims{};
  mh::Geo g{};
  mh::Dev d{};
  hipStream_t stream = nullptr;
  std::vector<void *> owned;                // every device allocation
  double scal[MH_NSCALARS] = {};
  bool scal_set[MH_NSCALARS] = {};
  bool started = false;
  // loop state (mf:3674-3714 and the per-step bookkeeping)
  int nstep = 0, ifend = 0;
  int hstep = 0;            // steps the host has queued (the device may stop earlier)
  std::map<std::string, hipGraphExec_t> graphs;  // captured step chunks, by step pattern
  int *ctl_host = nullptr;  // pinned copy of the device's CTL, for polling the end flag
  hipEvent_t ev_poll = nullptr;
  hipEvent_t ev_ahead[8] = {};  // completion of the last kAhead queued chunks
  long long nchunk = 0;         // chunks queued since the stream was last drained
  int nout[5] = {};       // printout steps (read from the device at mh_start)
  bool stop_req = false;  // mh_request_stop: end at the next step divisible by 10
  int device = 0;             // the HIP device the context was created on
  int scat_tk = 8, smooth_nw = 8;  // tile shapes in use (autotune at mh_start)
  bool te_sep = false;             // some row has IF_CUSP = 2 (the separation force runs)
  int *eff_tab = nullptr;        // mh_eficool's plane table and records (device)
  double *eff_out = nullptr;
  float *xl_scratch = nullptr;  // that device's address of mh_set_xlength_noalloc's scratch
  float DAMP = 0;                            // host mirror of the device's ramped DAMP (launch choices)
  bool mixed = false;                       // the SSS histories stored in BF16 (set at mh_start)
  unsigned short *h16[6] = {};              // their BF16 copies: DRO, DROVX, DROVR, DRORVT, DROE, DEL_DYNVIS
  unsigned short *store16 = nullptr;        // the mixed form's BF16 STORE (lane 0; the lanes hold their own)
  float PREFF = 0, ROLIM = 0, ROMIN = 0;
  float SFXIN_s = 0, SFTIN_s = 0;           // scaled by CFL/0.5 at the start of LOOP
  mh_resid last{};
  std::string err;
  // batched time steps (fast mode): the second and third variables' scratch
  int nlanes = 0;
  mh::Lane lanes[mh::kLanes];
};
