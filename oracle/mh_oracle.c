/* TEST INFRASTRUCTURE ONLY (the oracle).  Never linked into the product.
 *
 * Sequential FP32 restatement of the reference's time-marching loop; see
 * mh_oracle.h.  Every routine cites the reference lines it restates
 * (multall-open-18.3.f, "mf:" below).  Index macros take the reference's
 * 1-based (I,J,K) so loop bounds read as in the reference; all literals are
 * float so the arithmetic stays in single precision, statement by
 * statement, as the reference's does (build with -ffp-contract=off).
 *
 * Supported options (others are rejected by orc_start): ideal gas
 * (IFGAS=0), ITIMST=3 (scree; SSS and residual smoothing through F1/F2/F3
 * and NRSMTH), IM>2, KM>2, no cooling/bleed/shroud,
 * IF_REPEAT=0, no exit throttle or plate loss, IF_CUSP=0, IF_RESTART=0,
 * ILOS=0 (inviscid), ILOS=10 (mixing-length LOSS), 100 <= ILOS < 200
 * (NEW_LOSS) or ILOS >= 200 (SPAL_LOSS), each with the log-law, fixed-y+ or
 * Shih wall functions (YPLUSWALL).
 */
#include "mh_oracle.h"

#include <math.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

/* ----------------------------------------------------------------------- */
/* state                                                                    */

#define F3D(X) X(RO) X(ROVX) X(ROVR) X(RORVT) X(ROE) X(ROVT) X(ROWT) X(P) X(HO) X(PEFF) \
  X(VX) X(VR) X(VT) X(WT) X(T_STATIC) X(DRO) X(DROVX) X(DROVR) X(DRORVT) X(DROE)          \
  X(AQX) X(AQR) X(ABX) X(ABR) X(ASX) X(ASR) X(VOL) X(VOLOR) X(STEP) X(BSTEP) X(RSTEP)     \
  X(XFORCE) X(RFORCE) X(TFORCE) X(QSOURCE) X(FLOWX) X(FLOWT) X(FLOWR) X(XFLUX) X(TFLUX)  \
  X(RFLUX) X(SOURCE) X(STORE) X(TEMP1) X(SGEN) X(DWALLSQ) X(Y_PLUS) X(TRANS_DYN_VIS) X(DEL_DYNVIS)      \
  X(TRANS_KVIS) X(T_SOURCE) X(ROAVG_CELL) X(VISC_LAM) X(VISC_RAT) X(VISC_TURB) X(XLENGTH) X(FST_RAT)      \
  X(DYN_VIS_ML) X(ABS_VORT) X(TRANS_VORT) X(GRAD_TVIS) X(TVISGRADX) X(TVISGRADR) X(TVISGRADT) X(DPDS_CELL)
#define FJK(X) X(R) X(X) X(RAVG_CELL) X(ABT) X(WRABT) X(UBLADE) X(DX) X(DR) X(DS) X(SMERID) X(YPLUS_I1) X(YPLUS_IM)
/* read only for rows with IF_CUSP = 2: node R*THETA, blade thickness (J,K) */
#define FOPT(X) X(RTHETA) X(RT_THICK)
#define FIJ(X) X(YPLUS_K1) X(YPLUS_KM)
#define FIK(X) X(FTL) X(FBL) X(FTR) X(FBR)
#define F1D(X) X(FP) X(FU) X(FD) X(FR) X(FKU) X(FKD) X(PO1) X(TO1) X(RPO1) X(BSSIN) X(BSCOS) \
  X(BRSIN) X(BRCOS) X(VTIN) X(PD) X(PLIM) X(VM1) X(FACUP) X(FACDWN) X(WRAD) X(WHUB)         \
  X(WTIP) X(VLIMIT) X(ROLIMIT) X(AVG_CHG) X(AVG_BLK) X(STEP1) X(STEP2) X(STEPSBK)           \
  X(XLLIM_I1) X(XLLIM_IM) X(XLLIM_K1) X(XLLIM_KM) X(XLLIM_DWN) X(XLLIM_UP) X(CHORD)         \
  X(ROUGH_H) X(ROUGH_T) X(ROUGH_L) X(ROUGH_U) X(FSTURB) X(TURBVIS_DAMP) X(FILAM) X(FITURB) X(FIWAKE) X(FIUP) X(FKLAM) X(FKTURB) X(VISCOSY)  \
  X(REYNOLDS) X(ROVEXIT) X(SEP_THIK) X(SEP_DRAG) X(BS) X(BR)
#define I1D(X) X(IB1) X(IB2) X(KB1) X(KB2) X(NBLADE) X(IND) X(INDLE) X(INDMIX) X(INDTE)   \
  X(NROW) X(JSBLK) X(JB1) X(JB2) X(JSTART) X(JMIX) X(JLE) X(JTE) X(JLED) X(KTIPS) X(KTIPE) \
  X(KS1) X(KS2) X(NOUT) X(JTRAN_I1) X(JTRAN_IM) X(JTRAN_K1) X(JTRAN_KM) X(IF_CUSP) X(NUP_I1) X(NUP_IM) \
  X(N_WAKE)

#define SCAL(X) X(CFL) X(DAMP) X(DAMPIN) X(MACHLIM) X(F_PDOWN) X(CONLIM) X(SFXIN) X(SFTIN)    \
  X(SFX) X(SFT) X(FAC_4TH) X(F1) X(F2) X(F3) X(RSMTH) X(SFEXIT) X(FSMTHB) X(RFIN) X(RFIN1)    \
  X(RFIN2) X(PDOWN_HUB) X(PDOWN_TIP) X(FEXTRAP) X(RFMIX) X(FANGLE) X(CP) X(GA) X(RGAS) X(CV)  \
  X(FGA) X(GA1) X(RCP) X(VSOUND) X(FACMIXUP) X(FMIXUP) X(REYNO) X(RF_VIS) X(FTRANS)           \
  X(PRANDTL) X(YPLUSWALL) X(TURBVIS_LIM) X(VISLAM) X(FTCOND) X(PREFF) X(ROLIM) X(ROMIN)       \
  X(RO_REF) X(EAVG) X(EMAX) X(ECONT) X(VRMS) X(VMAX) X(VREF) X(TCOOL_MIN) X(SFEX1) X(SFEXH)   \
  X(FLOWRAT) X(YPLAM) X(YPTURB) X(CFWALL) X(FAC_STMIX) X(FAC_ST0) X(FAC_ST1) X(FAC_ST2)       \
  X(FAC_ST3) X(FAC_SFVIS) X(FAC_VORT) X(FAC_PGRAD) X(VORT_MAX) X(FLOWIN) X(RFLOW) X(RFINBC) X(PO_IN_MID)   \
  X(VT_IN_MID)
#define ISCAL(X) X(IM) X(JM) X(KM) X(NROWS) X(IMID) X(KMID) X(NSTEP) X(NSTEPS_MAX) X(NCHANGE) \
  X(NRSMTH) X(NSFEXIT) X(IN_PRESS) X(IN_VTAN) X(IN_VR) X(IPOUT) X(IFMIX) X(IR) X(JR) X(KR)    \
  X(ILOS) X(NLOS) X(NMIXUP) X(NIB1) X(NJB1) X(NKB1) X(NIB2) X(NJB2) X(NKB2) X(NSBLK) X(IFEND) \
  X(IFPRINT) X(ITIMST) X(IF_RESTART) X(IFGAS) X(IN_FLOW) X(IF_REPEAT) X(IBOUND) X(NINMOD)

struct orc {
#define X(n) float *n;
  F3D(X) FJK(X) FIK(X) FIJ(X) F1D(X) FOPT(X)
#undef X
#define X(n) int *n;
  I1D(X)
#undef X
#define X(n) float n;
  SCAL(X)
#undef X
#define X(n) int n;
  ISCAL(X)
#undef X
  float TCOND;
  float *CHECK_FLOW, *BLADE_FLOW; /* LOOP's SAVEd locals (mf:3672-3673) */
  float *B1CHG, *B2CHG, *SBCHG;   /* TSTEP's block accumulators (mf:5482) */
};

static void orc_repeating_stage(orc_t *o);

orc_t *orc_create(void) { return (orc_t *)calloc(1, sizeof(orc_t)); }
void orc_destroy(orc_t *o) {
  if (!o) return;
  free(o->CHECK_FLOW); free(o->BLADE_FLOW); free(o->B1CHG); free(o->B2CHG); free(o->SBCHG);
  free(o);
}

int orc_bind(orc_t *o, const char *name, void *ptr) {
#define X(n) if (!strcmp(name, #n)) { o->n = (float *)ptr; return 0; }
  F3D(X) FJK(X) FIK(X) FIJ(X) F1D(X) FOPT(X)
#undef X
#define X(n) if (!strcmp(name, #n)) { o->n = (int *)ptr; return 0; }
  I1D(X)
#undef X
  return -1;
}

int orc_scalar(orc_t *o, const char *name, double v) {
#define X(n) if (!strcmp(name, #n)) { o->n = (float)v; return 0; }
  SCAL(X)
#undef X
#define X(n) if (!strcmp(name, #n)) { o->n = (int)v; return 0; }
  ISCAL(X)
#undef X
  return -1;
}

double orc_get_scalar(orc_t *o, const char *name) {
#define X(n) if (!strcmp(name, #n)) return (double)o->n;
  SCAL(X) ISCAL(X)
#undef X
  return NAN;
}

/* 1-based index helpers, (I,J,K) as in the reference */
#define IM_ (o->IM)
#define JM_ (o->JM)
#define KM_ (o->KM)
#define A3(a, i, j, k) (o->a[((i) - 1) + (size_t)IM_ * (((j) - 1) + (size_t)JM_ * ((k) - 1))])
#define D3(p, i, j, k) ((p)[((i) - 1) + (size_t)IM_ * (((j) - 1) + (size_t)JM_ * ((k) - 1))])
#define AJK(a, j, k) (o->a[((j) - 1) + (size_t)JM_ * ((k) - 1)])
#define AIK(a, i, k) (o->a[((i) - 1) + (size_t)IM_ * ((k) - 1)])
#define AIJ(a, i, j) (o->a[((i) - 1) + (size_t)IM_ * ((j) - 1)])
#define A1(a, n) (o->a[(n) - 1])
#define NRSZ 21 /* leading dimension of the (NRS,MAXKI) LOSS tables, mf:commall :73-75 */
#define ANR(a, nr, i) (o->a[((nr) - 1) + (size_t)NRSZ * ((i) - 1)])
#define STEP1_(a, b, c) (o->STEP1[((a) - 1) + (size_t)o->NIB1 * (((b) - 1) + (size_t)(o->NJB1 + 1) * ((c) - 1))])
#define STEP2_(a, b, c) (o->STEP2[((a) - 1) + (size_t)o->NIB2 * (((b) - 1) + (size_t)(o->NJB2 + 1) * ((c) - 1))])
#define B1_(a, b, c) (o->B1CHG[((a) - 1) + (size_t)o->NIB1 * (((b) - 1) + (size_t)(o->NJB1 + 1) * ((c) - 1))])
#define B2_(a, b, c) (o->B2CHG[((a) - 1) + (size_t)o->NIB2 * (((b) - 1) + (size_t)(o->NJB2 + 1) * ((c) - 1))])

/* ----------------------------------------------------------------------- */
/* set-up at the top of LOOP (mf:3674-3714)                                 */

int orc_start(orc_t *o) {
#define X(n) if (!o->n) return -100;
  F3D(X) FJK(X) FIK(X) FIJ(X)
#undef X
  if (o->IFGAS != 0 || o->ITIMST != 3 || o->IM <= 2 || o->KM <= 2 || o->IF_RESTART != 0 ||
      (o->IF_REPEAT != 0 && (o->NINMOD < 1 || !o->BS || !o->BR)))
    return -2;
  if (o->ILOS != 0 && o->ILOS != 10 && o->ILOS < 100) return -3;
  for (int NR = 1; o->IF_CUSP && NR <= o->NROWS; NR++)
    if (o->IF_CUSP[NR - 1] == 2 && (!o->RTHETA || !o->RT_THICK || !o->NUP_I1 || !o->NUP_IM || !o->N_WAKE ||
                                    !o->SEP_THIK || !o->SEP_DRAG))
      return -100;
  const int IM = o->IM, JM = o->JM, KM = o->KM, IMID = o->IMID, KMID = o->KMID;
  (void)IM;
  o->RO_REF = A1(PO1, KMID) / (o->RGAS * A1(TO1, KMID));
  o->ROMIN = fminf(A3(RO, IMID, 1, KMID), A3(RO, IMID, JM, KMID));
  o->ROLIM = 0.25f * o->ROMIN;
  for (int K = 1; K <= KM; K++) A1(PLIM, K) = 0.9f * A1(PO1, K) + 0.1f * A3(P, 1, 1, K);
  o->PREFF = 0.5f * (A3(P, IMID, 1, KMID) + A3(P, IMID, JM, KMID));
  o->SFXIN = o->SFXIN * o->CFL / 0.5f;
  o->SFTIN = o->SFTIN * o->CFL / 0.5f;
  o->RFIN1 = 1.0f - o->RFIN;
  o->RFIN2 = 2.f * o->RFIN;
  o->ECONT = 1.f;
  o->EAVG = 1.0e06f;
  o->EMAX = 1.0e06f;
  o->TCOOL_MIN = 1.0e6f;
  o->IFEND = 0;
  o->NSTEP = 0;
  o->SFEX1 = 1.f - o->SFEXIT;
  o->SFEXH = 0.5f * o->SFEXIT;
  free(o->CHECK_FLOW); free(o->BLADE_FLOW); free(o->B1CHG); free(o->B2CHG); free(o->SBCHG);
  o->CHECK_FLOW = (float *)calloc(JM + 1, sizeof(float));
  o->BLADE_FLOW = (float *)calloc(JM + 1, sizeof(float));
  o->B1CHG = (float *)calloc((size_t)o->NIB1 * (o->NJB1 + 1) * o->NKB1 + 1, sizeof(float));
  o->B2CHG = (float *)calloc((size_t)o->NIB2 * (o->NJB2 + 1) * o->NKB2 + 1, sizeof(float));
  o->SBCHG = (float *)calloc(o->NSBLK + 1, sizeof(float));
  return 0;
}

/* ----------------------------------------------------------------------- */
/* primitive variables, mf:3826-3897                                        */

void orc_primitives(orc_t *o) {
  const int IM = o->IM, JM = o->JM, KM = o->KM, JMM1 = JM - 1;
  for (int K = 1; K <= KM; K++)
    for (int J = 2; J <= JM; J++) {
      float RRNOW = 1 / AJK(R, J, K);
      for (int I = 1; I <= IM; I++) {
        float RECIP = 1.0f / A3(RO, I, J, K);
        A3(VX, I, J, K) = A3(ROVX, I, J, K) * RECIP;
        A3(VR, I, J, K) = A3(ROVR, I, J, K) * RECIP;
        A3(ROVT, I, J, K) = A3(RORVT, I, J, K) * RRNOW;
        A3(VT, I, J, K) = A3(ROVT, I, J, K) * RECIP;
      }
    }
  if (o->NSTEP % 5 == 0 || o->IFEND == 1) {
    float VRMS = 0.0f, VMAX = 0.0f;
    for (int K = 1; K <= KM; K++)
      for (int J = 2; J <= JM; J++)
        for (int I = 1; I <= IM; I++) {
          float EKE = A3(VR, I, J, K) * A3(VR, I, J, K) + A3(WT, I, J, K) * A3(WT, I, J, K) +
                      A3(VX, I, J, K) * A3(VX, I, J, K);
          if (EKE > VMAX) VMAX = EKE;
          VRMS = VRMS + EKE;
        }
    o->VMAX = sqrtf(VMAX);
    o->VRMS = VRMS;
  }
  for (int K = 1; K <= KM; K++)
    for (int J = 2; J <= JMM1; J++)
      for (int I = 1; I <= IM; I++) {
        float RONOW = A3(RO, I, J, K);
        float EKE = .5f * (A3(VR, I, J, K) * A3(VR, I, J, K) + A3(VT, I, J, K) * A3(VT, I, J, K) +
                           A3(VX, I, J, K) * A3(VX, I, J, K));
        float TSTATIC = (A3(ROE, I, J, K) / RONOW - EKE) / o->CV;
        A3(P, I, J, K) = RONOW * o->RGAS * TSTATIC;
        A3(HO, I, J, K) = o->CP * TSTATIC + EKE;
        A3(T_STATIC, I, J, K) = TSTATIC;
      }
}

/* ----------------------------------------------------------------------- */
/* SMOOTH_RESID, mf:14833-14917                                             */

static void smooth_resid(orc_t *o, float *var, float SMFAC, int NSMTH) {
  const int IM = o->IM, KM = o->KM, IMM1 = IM - 1, IMM2 = IM - 2, KMM1 = KM - 1, KMM2 = KM - 2;
  float *TEMPVAR = (float *)malloc(sizeof(float) * (size_t)(o->JM + IM + KM + 8));
  float SFAC1 = 1.0f - SMFAC, SFACH = 0.5f * SMFAC;
  for (int NSM = 1; NSM <= NSMTH; NSM++) {
    for (int NR = 1; NR <= o->NROWS; NR++) {
      int J1 = A1(JSTART, NR), J1P1 = J1 + 1, J1P2 = J1 + 2, J2 = A1(JMIX, NR), J2M1 = J2 - 1;
      for (int K = 1; K <= KMM1; K++)
        for (int I = 1; I <= IMM1; I++) {
          for (int J = J1P2; J <= J2M1; J++)
            TEMPVAR[J] = SFAC1 * D3(var, I, J, K) + SFACH * (D3(var, I, J + 1, K) + D3(var, I, J - 1, K));
          TEMPVAR[J1P1] = D3(var, I, J1P1, K) + SFACH * (D3(var, I, J1P2, K) - D3(var, I, J1P1, K));
          TEMPVAR[J2] = D3(var, I, J2, K) + SFACH * (D3(var, I, J2M1, K) - D3(var, I, J2, K));
          for (int J = J1P1; J <= J2; J++) D3(var, I, J, K) = TEMPVAR[J];
        }
      for (int K = 1; K <= KMM1; K++)
        for (int J = J1P1; J <= J2; J++) {
          for (int I = 2; I <= IMM2; I++)
            TEMPVAR[I] = SFAC1 * D3(var, I, J, K) + SFACH * (D3(var, I + 1, J, K) + D3(var, I - 1, J, K));
          TEMPVAR[1] = D3(var, 1, J, K) + SFACH * (D3(var, 2, J, K) - D3(var, 1, J, K));
          TEMPVAR[IMM1] = D3(var, IMM1, J, K) + SFACH * (D3(var, IMM2, J, K) - D3(var, IMM1, J, K));
          for (int I = 1; I <= IMM1; I++) D3(var, I, J, K) = TEMPVAR[I];
        }
      for (int I = 1; I <= IMM1; I++)
        for (int J = J1P1; J <= J2; J++) {
          for (int K = 2; K <= KMM2; K++)
            TEMPVAR[K] = SFAC1 * D3(var, I, J, K) + SFACH * (D3(var, I, J, K + 1) + D3(var, I, J, K - 1));
          TEMPVAR[1] = D3(var, I, J, 1) + SFACH * (D3(var, I, J, 2) - D3(var, I, J, 1));
          TEMPVAR[KMM1] = D3(var, I, J, KMM1) + SFACH * (D3(var, I, J, KMM2) - D3(var, I, J, KMM1));
          for (int K = 1; K <= KMM1; K++) D3(var, I, J, K) = TEMPVAR[K];
        }
    }
  }
  free(TEMPVAR);
}

/* ----------------------------------------------------------------------- */
/* SMOOTH_VAR, mf:15352-15594                                               */

void orc_smooth_var(orc_t *o, float *D) {
  const int IM = o->IM, JM = o->JM, KM = o->KM, IMM1 = IM - 1, IMM2 = IM - 2, KMM1 = KM - 1,
            KMM2 = KM - 2;
  const int NUM3 = 3;
  const float SFX = o->SFX, SFX1 = 1.f - SFX, FAC_4TH = o->FAC_4TH;
  const float SFXB = o->FSMTHB * SFX, SFXB1 = 1.0f - SFXB, SFXBH = 0.5f * SFXB;
  const float SFXH = 0.5f * SFX, SFXHM1 = 1.f - SFXH, SFX14 = 0.25f * SFX;
  const float SFT = o->SFT, SFT1 = 1.0f - SFT, SFTH = 0.5f * SFT;
  int nmax = (IM > KM ? IM : KM) + 2;
  float *AVG = (float *)malloc(sizeof(float) * nmax), *CURVE = (float *)malloc(sizeof(float) * nmax),
        *SCURVE = (float *)malloc(sizeof(float) * nmax);
  float *STORE = o->STORE;
  if (!(SFX < 0.0001f)) {
    for (int NR = 1; NR <= o->NROWS; NR++) {
      int J1 = A1(JSTART, NR), J2 = A1(JMIX, NR);
      for (int J = J1 + 3; J <= J2 - 3; J++)
        for (int K = 1; K <= KM; K++)
          for (int I = 1; I <= IM; I++) {
            float AVERG = 0.5f * (D3(D, I, J + 1, K) + D3(D, I, J - 1, K));
            float CURV = AVERG - 0.5f * D3(D, I, J, K) - 0.25f * (D3(D, I, J - 2, K) + D3(D, I, J + 2, K));
            D3(STORE, I, J, K) = SFX1 * D3(D, I, J, K) + SFX * (AVERG + FAC_4TH * CURV);
          }
      for (int K = 1; K <= KM; K++)
        for (int I = 1; I <= IM; I++) {
          D3(STORE, I, J2, K) = D3(D, I, J2, K);
          D3(STORE, I, J1, K) = D3(D, I, J1, K);
        }
      for (int K = 1; K <= KM; K++)
        for (int I = 1; I <= IM; I++) {
          D3(STORE, I, J1 + 2, K) = SFXB1 * D3(D, I, J1 + 2, K) + SFXBH * (D3(D, I, J1 + 3, K) + D3(D, I, J1 + 1, K));
          D3(STORE, I, J1 + 1, K) = SFXB1 * D3(D, I, J1 + 1, K) +
                                    SFXB * (1.25f * D3(D, I, J1 + 2, K) + 0.5f * D3(D, I, J1 + 3, K) -
                                            0.75f * D3(D, I, J1 + 4, K));
          D3(STORE, I, J2 - 2, K) = SFXB1 * D3(D, I, J2 - 2, K) + SFXBH * (D3(D, I, J2 - 1, K) + D3(D, I, J2 - 3, K));
          D3(STORE, I, J2 - 1, K) = SFXB1 * D3(D, I, J2 - 1, K) +
                                    SFXB * (1.25f * D3(D, I, J2 - 2, K) + 0.5f * D3(D, I, J2 - 3, K) -
                                            0.75f * D3(D, I, J2 - 4, K));
        }
    }
    for (int NR = 1; NR <= o->NROWS; NR++) {
      int JS = A1(JLED, NR) - 2, JE = JS + 1;
      for (int J = JS; J <= JE; J++)
        for (int K = 1; K <= KM; K++) {
          D3(STORE, 1, J, K) = SFXHM1 * D3(D, 1, J, K) + SFX14 * (D3(D, 2, J, K) + D3(D, IMM1, J, K));
          D3(STORE, IM, J, K) = SFXHM1 * D3(D, IM, J, K) + SFX14 * (D3(D, 2, J, K) + D3(D, IMM1, J, K));
        }
      int J = A1(JLED, NR);
      for (int K = 1; K <= KM; K++) {
        D3(STORE, 1, J, K) = SFXHM1 * D3(D, 1, J, K) + SFX14 * (D3(D, 1, J + 1, K) + D3(D, IM, J, K));
        D3(STORE, IM, J, K) = SFXHM1 * D3(D, IM, J, K) + SFX14 * (D3(D, IM, J + 1, K) + D3(D, 1, J, K));
      }
    }
    for (int NR = 1; NR <= o->NROWS; NR++) {
      int J1 = A1(JSTART, NR), J2 = A1(JMIX, NR);
      for (int J = J1; J <= J2; J++)
        for (int K = 1; K <= KM; K++)
          for (int I = 1; I <= IM; I++) D3(D, I, J, K) = D3(STORE, I, J, K);
    }
  }
  if (!(SFT < 0.0001f)) {
    for (int J = 2; J <= JM; J++) {
      for (int K = 1; K <= KM; K++) {
        for (int I = 2; I <= IMM1; I++) {
          AVG[I] = A1(FD, I) * D3(D, I + 1, J, K) + A1(FU, I) * D3(D, I - 1, J, K);
          CURVE[I] = D3(D, I, J, K) - AVG[I];
        }
        if (!(A1(IND, J) == 1 || A1(INDLE, J) == 1)) {
          AVG[1] = 0.5f * (D3(D, 2, J, K) + D3(D, IMM1, J, K));
          AVG[IM] = AVG[1];
          CURVE[1] = D3(D, 1, J, K) - AVG[1];
          CURVE[IM] = CURVE[1];
          SCURVE[1] = 0.5f * (CURVE[2] + CURVE[IMM1]);
          SCURVE[IM] = SCURVE[1];
        } else {
          AVG[1] = D3(D, 2, J, K) + A1(FU, 1) * (D3(D, 2, J, K) - D3(D, NUM3, J, K));
          AVG[IM] = D3(D, IMM1, J, K) + A1(FU, IM) * (D3(D, IMM1, J, K) - D3(D, IMM2, J, K));
          CURVE[1] = 0.0f;
          CURVE[IM] = 0.0f;
          SCURVE[1] = 0.0f;
          SCURVE[IM] = 0.0f;
        }
        for (int I = 2; I <= IMM1; I++) SCURVE[I] = A1(FU, I) * CURVE[I - 1] + A1(FD, I) * CURVE[I + 1];
        for (int I = 1; I <= IM; I++)
          D3(D, I, J, K) = SFT1 * D3(D, I, J, K) + SFT * (AVG[I] + FAC_4TH * SCURVE[I]);
      }
      for (int I = 1; I <= IM; I++) {
        for (int K = 2; K <= KMM1; K++) {
          AVG[K] = A1(FKD, K) * D3(D, I, J, K + 1) + A1(FKU, K) * D3(D, I, J, K - 1);
          CURVE[K] = D3(D, I, J, K) - AVG[K];
        }
        AVG[1] = D3(D, I, J, 2) + 0.5f * A1(FKU, 1) * (D3(D, I, J, 2) - D3(D, I, J, 3));
        AVG[KM] = D3(D, I, J, KMM1) + 0.5f * A1(FKU, KM) * (D3(D, I, J, KMM1) - D3(D, I, J, KMM2));
        CURVE[1] = 0.0f;
        CURVE[KM] = 0.0f;
        SCURVE[1] = 0.0f;
        SCURVE[KM] = 0.0f;
        for (int K = 2; K <= KMM1; K++) SCURVE[K] = A1(FKU, K) * CURVE[K - 1] + A1(FKD, K) * CURVE[K + 1];
        for (int K = 1; K <= KM; K++)
          D3(D, I, J, K) = SFT1 * D3(D, I, J, K) + SFT * (AVG[K] + FAC_4TH * SCURVE[K]);
      }
      D3(D, 1, J, 1) = SFT1 * D3(D, 1, J, 1) + SFTH * (D3(D, 2, J, 1) + D3(D, 1, J, 2));
      D3(D, IM, J, 1) = SFT1 * D3(D, IM, J, 1) + SFTH * (D3(D, IMM1, J, 1) + D3(D, IM, J, 2));
      D3(D, 1, J, KM) = SFT1 * D3(D, 1, J, KM) + SFTH * (D3(D, 1, J, KMM1) + D3(D, 2, J, KM));
      D3(D, IM, J, KM) = SFT1 * D3(D, IM, J, KM) + SFTH * (D3(D, IMM1, J, KM) + D3(D, IM, J, KMM1));
    }
  }
  if (!(o->SFEXIT < 0.0001f)) {
    int JSSTART = JM + 1 - o->NSFEXIT;
    float SFEX1 = o->SFEX1, SFEXH = o->SFEXH, SFEXIT = o->SFEXIT;
    for (int J = JSSTART; J <= JM; J++) {
      for (int K = 1; K <= KM; K++) {
        for (int I = 2; I <= IMM1; I++)
          D3(D, I, J, K) = SFEX1 * D3(D, I, J, K) + SFEXH * (D3(D, I + 1, J, K) + D3(D, I - 1, J, K));
        D3(D, 1, J, K) = D3(D, 1, J, K) + SFEXIT * (2.f * D3(D, 2, J, K) - D3(D, NUM3, J, K) - D3(D, 1, J, K));
        D3(D, IM, J, K) = D3(D, IM, J, K) + SFEXIT * (2.f * D3(D, IMM1, J, K) - D3(D, IMM2, J, K) - D3(D, IM, J, K));
      }
      for (int I = 1; I <= IM; I++) {
        for (int K = 2; K <= KMM1; K++)
          D3(D, I, J, K) = SFEX1 * D3(D, I, J, K) + SFEXH * (D3(D, I, J, K - 1) + D3(D, I, J, K + 1));
        D3(D, I, J, 1) = D3(D, I, J, 1) + SFEXIT * (2.f * D3(D, I, J, 2) - D3(D, I, J, 3) - D3(D, I, J, 1));
        D3(D, I, J, KM) = D3(D, I, J, KM) + SFEXIT * (2.f * D3(D, I, J, KMM1) - D3(D, I, J, KMM2) - D3(D, I, J, KM));
      }
    }
  }
  free(AVG); free(CURVE); free(SCURVE);
}

/* ----------------------------------------------------------------------- */
/* TSTEP, mf:5470-5932                                                      */

void orc_tstep(orc_t *o, float *D, float *DIFF, int NCALL) {
  const int IM = o->IM, JM = o->JM, KM = o->KM, IMM1 = IM - 1, KMM1 = KM - 1, IMID = o->IMID;
  const int NRWSM1 = o->NROWS - 1;
  float *STORE = o->STORE;
  for (int K = 1; K <= o->NKB1; K++)
    for (int J = 1; J <= o->NJB1 + 1; J++)
      for (int I = 1; I <= o->NIB1; I++) B1_(I, J, K) = 0.0f;
  for (int K = 1; K <= o->NKB2; K++)
    for (int J = 1; J <= o->NJB2 + 1; J++)
      for (int I = 1; I <= o->NIB2; I++) B2_(I, J, K) = 0.0f;
  for (int J = 1; J <= o->NSBLK; J++) o->SBCHG[J] = 0.0f;
  /* tip-gap flux balance, mf:5503-5510 */
  for (int J = 2; J <= JM; J++) {
    int NR = A1(NROW, J);
    if (A1(KTIPS, NR) <= 0) continue;
    for (int K = A1(KTIPS, NR); K <= A1(KTIPE, NR) - 1; K++) {
      A3(TFLUX, 1, J, K) = 0.5f * (A3(TFLUX, 1, J, K) + A3(TFLUX, IM, J, K));
      A3(TFLUX, IM, J, K) = A3(TFLUX, 1, J, K);
    }
  }
  /* upstream-face flux extrapolation at mixing planes, mf:5516-5555 */
  if (!(o->FEXTRAP < 0.001f)) {
    for (int NR = 1; NR <= NRWSM1; NR++) {
      int J = A1(JMIX, NR), JP1 = J + 1, JM1 = J - 1, JP2 = J + 2;
      for (int K = 1; K <= KMM1; K++) {
        float FLOWDIRN = -A3(FLOWX, IMID, J, K);
        float AVFLXJM1 = 0.0f, AVFLXJP2 = 0.0f;
        for (int I = 1; I <= IMM1; I++) {
          AVFLXJM1 = AVFLXJM1 + A3(XFLUX, I, JM1, K);
          AVFLXJP2 = AVFLXJP2 + A3(XFLUX, I, JP2, K);
        }
        for (int I = 1; I <= IMM1; I++) {
          if (FLOWDIRN > 0.0f) {
            float DFLUXJM1 = A3(XFLUX, I, JM1, K) - AVFLXJM1 * A1(FP, I);
            float DFLUX = DFLUXJM1 * o->FEXTRAP;
            A3(SOURCE, I, J, K) = A3(SOURCE, I, J, K) - DFLUX;
          } else {
            float DFLUXJP2 = A3(XFLUX, I, JP2, K) - AVFLXJP2 * A1(FP, I);
            float DFLUX = DFLUXJP2 * o->FEXTRAP;
            A3(SOURCE, I, JP2, K) = A3(SOURCE, I, JP2, K) + DFLUX;
          }
          A3(SOURCE, I, JP1, K) = 0.0f;
        }
      }
    }
  }
  /* residual and scree/SSS history, mf:5560-5570 */
  for (int K = 1; K <= KMM1; K++)
    for (int J = 2; J <= JM; J++) {
      float RATPITCH = (float)A1(NBLADE, J - 1) / (float)A1(NBLADE, J);
      for (int I = 1; I <= IMM1; I++) {
        float DELTA = A3(XFLUX, I, J, K) - A3(XFLUX, I, J - 1, K) * RATPITCH + A3(RFLUX, I, J, K) -
                      A3(RFLUX, I, J, K + 1) + A3(TFLUX, I, J, K) - A3(TFLUX, I + 1, J, K) - A3(SOURCE, I, J, K);
        D3(STORE, I, J, K) = o->F1 * DELTA + o->F2 * D3(DIFF, I, J, K);
        D3(DIFF, I, J, K) = DELTA + o->F3 * D3(DIFF, I, J, K);
      }
    }
  /* pitchwise averaging of the change at mixing planes, mf:5578-5607 */
  for (int NR = 1; NR <= NRWSM1; NR++) {
    int J = A1(JMIX, NR);
    for (int K = 1; K <= KMM1; K++) {
      float FLOWDIRN = -A3(FLOWX, IMID, J, K);
      if ((FLOWDIRN > 0.0f) || (o->FEXTRAP < 0.001f)) {
        int JMX = J + 2;
        float SUM_STORE = 0.0f;
        for (int I = 1; I <= IMM1; I++) SUM_STORE = SUM_STORE + D3(STORE, I, JMX, K);
        for (int I = 1; I <= IMM1; I++) D3(STORE, I, JMX, K) = SUM_STORE * A1(FP, I);
      }
      if ((FLOWDIRN < 0.0f) || (o->FEXTRAP < 0.001f)) {
        float SUM_STORE = 0.0f;
        for (int I = 1; I <= IMM1; I++) SUM_STORE = SUM_STORE + D3(STORE, I, J, K);
        for (int I = 1; I <= IMM1; I++) D3(STORE, I, J, K) = SUM_STORE * A1(FP, I);
      }
    }
  }
  /* multigrid block sums and update, mf:5612-5660 */
  int mg = !(o->IR <= 1 && o->JR <= 1 && o->KR <= 1);
  if (mg) {
    for (int K = 1; K <= KMM1; K++) {
      int K1 = A1(KB1, K), K2 = A1(KB2, K);
      for (int J = 2; J <= JM; J++) {
        int JSB = A1(JSBLK, J), J1 = A1(JB1, J), J2 = A1(JB2, J);
        for (int I = 1; I <= IMM1; I++) {
          int I1 = A1(IB1, I), I2 = A1(IB2, I);
          float DELTA = D3(STORE, I, J, K);
          B1_(I1, J1, K1) = B1_(I1, J1, K1) + DELTA;
          B2_(I2, J2, K2) = B2_(I2, J2, K2) + DELTA;
          o->SBCHG[JSB] = o->SBCHG[JSB] + DELTA;
        }
      }
    }
  }
  for (int K = 1; K <= KMM1; K++) {
    int K1 = A1(KB1, K), K2 = A1(KB2, K);
    for (int J = 2; J <= JM; J++) {
      int JSB = A1(JSBLK, J), J1 = A1(JB1, J), J2 = A1(JB2, J);
      for (int I = 1; I <= IMM1; I++) {
        int I1 = A1(IB1, I), I2 = A1(IB2, I);
        float DELTA = D3(STORE, I, J, K) * A3(STEP, I, J, K) +
                      (B1_(I1, J1, K1) * STEP1_(I1, J1, K1) + B2_(I2, J2, K2) * STEP2_(I2, J2, K2) +
                       o->SBCHG[JSB] * A1(STEPSBK, JSB)) *
                          A3(RSTEP, I, J, K);
        D3(STORE, I, J, K) = DELTA;
      }
    }
  }
  if (o->NRSMTH > 0) smooth_resid(o, STORE, o->RSMTH, o->NRSMTH);
  /* negative feedback, mf:5677-5724 */
  if (!(o->DAMP < 2.0f || o->DAMP > 100)) {
    for (int NR = 1; NR <= o->NROWS; NR++) {
      float SUMCHG = 0.0f;
      int JST = A1(JSTART, NR) + 1, JEN = A1(JMIX, NR) - 1;
      int JCHANGE = JEN - JST + 1;
      int NSUM = IMM1 * KMM1 * JCHANGE;
      for (int K = 1; K <= KMM1; K++)
        for (int J = JST; J <= JEN; J++)
          for (int I = 1; I <= IMM1; I++) SUMCHG = SUMCHG + fabsf(D3(STORE, I, J, K));
      A1(AVG_CHG, NR) = SUMCHG / (float)NSUM;
    }
    int NROWS = o->NROWS;
    if (NROWS == 1) A1(AVG_BLK, 1) = A1(AVG_CHG, 1);
    if (NROWS == 2) {
      A1(AVG_BLK, 2) = 0.5f * (A1(AVG_CHG, 1) + A1(AVG_CHG, 2));
      A1(AVG_BLK, 1) = A1(AVG_BLK, 2);
    }
    if (NROWS > 2) {
      for (int NR = 2; NR <= NROWS - 1; NR++)
        A1(AVG_BLK, NR) = 0.25f * (A1(AVG_CHG, NR - 1) + A1(AVG_CHG, NR + 1)) + 0.5f * A1(AVG_CHG, NR);
      A1(AVG_BLK, NROWS) = 0.5f * (A1(AVG_CHG, NROWS - 1) + A1(AVG_CHG, NROWS));
      A1(AVG_BLK, 1) = 0.5f * (A1(AVG_CHG, 1) + A1(AVG_CHG, 2));
    }
    for (int K = 1; K <= KMM1; K++)
      for (int J = 2; J <= JM; J++) {
        int NR = A1(NROW, J);
        for (int I = 1; I <= IMM1; I++) {
          float DELTA = D3(STORE, I, J, K);
          float ABSCHG = fabsf(DELTA);
          float FDAMP = ABSCHG / A1(AVG_BLK, NR);
          D3(STORE, I, J, K) = DELTA / (1.f + FDAMP / o->DAMP);
        }
      }
  }
  int ILAST = IM;
  if (NCALL == 2) {
    /* the SA variable is cell-centred: add the change, carry the pitch
     * average across each mixing plane and smooth, mf:5736-5777 */
    for (int K = 1; K <= KMM1; K++)
      for (int J = 2; J <= JM; J++)
        for (int I = 1; I <= IMM1; I++) D3(D, I, J, K) = D3(D, I, J, K) + D3(STORE, I, J, K);
    for (int NR = 1; NR <= o->NROWS - 1; NR++) {
      int J = A1(JMIX, NR);
      float FMULT = A1(TURBVIS_DAMP, NR);
      for (int K = 1; K <= KMM1; K++) {
        float AVG = 0.0f;
        for (int I = 1; I <= IMM1; I++) AVG = AVG + A1(FP, I) * D3(D, I, J, K);
        for (int JAV = J + 1; JAV <= J + 2; JAV++)
          for (int I = 1; I <= IMM1; I++) D3(D, I, JAV, K) = FMULT * AVG;
      }
    }
    smooth_resid(o, D, o->FAC_SFVIS * o->SFT, 1);
    ILAST = IMM1;
    goto bcs;
  }
  /* second pitchwise averaging at mixing planes, mf:5786-5818 */
  for (int NR = 1; NR <= NRWSM1; NR++) {
    int J = A1(JMIX, NR);
    for (int K = 1; K <= KMM1; K++) {
      float FLOWDIRN = -A3(FLOWX, IMID, J, K);
      if ((FLOWDIRN > 0.0f) || (o->FEXTRAP < 0.001f)) {
        int JMX = J + 2;
        float SUM_STORE = 0.0f;
        for (int I = 1; I <= IMM1; I++) SUM_STORE = SUM_STORE + D3(STORE, I, JMX, K);
        SUM_STORE = SUM_STORE / (float)IMM1;
        for (int I = 1; I <= IMM1; I++) D3(STORE, I, JMX, K) = SUM_STORE;
      }
      if ((FLOWDIRN < 0.0f) || (o->FEXTRAP < 0.001f)) {
        float SUM_STORE = 0.0f;
        for (int I = 1; I <= IMM1; I++) SUM_STORE = SUM_STORE + D3(STORE, I, J, K);
        SUM_STORE = SUM_STORE / (float)IMM1;
        for (int I = 1; I <= IMM1; I++) D3(STORE, I, J, K) = SUM_STORE;
      }
    }
  }
  /* distribute the cell change to its 8 corners, mf:5828-5842 */
  for (int K = 1; K <= KMM1; K++)
    for (int J = 2; J <= JM; J++)
      for (int I = 1; I <= IMM1; I++) {
        float ADD = D3(STORE, I, J, K);
        D3(D, I, J, K) = D3(D, I, J, K) + ADD * AIK(FBL, I, K) * A1(FACDWN, J);
        D3(D, I + 1, J, K) = D3(D, I + 1, J, K) + ADD * AIK(FBR, I, K) * A1(FACDWN, J);
        D3(D, I, J, K + 1) = D3(D, I, J, K + 1) + ADD * AIK(FTL, I, K) * A1(FACDWN, J);
        D3(D, I + 1, J, K + 1) = D3(D, I + 1, J, K + 1) + ADD * AIK(FTR, I, K) * A1(FACDWN, J);
        D3(D, I, J - 1, K) = D3(D, I, J - 1, K) + ADD * AIK(FBL, I, K) * A1(FACUP, J);
        D3(D, I + 1, J - 1, K) = D3(D, I + 1, J - 1, K) + ADD * AIK(FBR, I, K) * A1(FACUP, J);
        D3(D, I, J - 1, K + 1) = D3(D, I, J - 1, K + 1) + ADD * AIK(FTL, I, K) * A1(FACUP, J);
        D3(D, I + 1, J - 1, K + 1) = D3(D, I + 1, J - 1, K + 1) + ADD * AIK(FTR, I, K) * A1(FACUP, J);
      }
  orc_smooth_var(o, D);
bcs:
  /* periodicity, tip gap and mixing-plane averages of D, mf:5858-5926 */
  for (int J = 1; J <= JM; J++) {
    if ((A1(IND, J) == 1) || (A1(INDLE, J) == 1)) continue;
    for (int K = 1; K <= KM; K++) {
      D3(D, 1, J, K) = 0.5f * (D3(D, 1, J, K) + D3(D, ILAST, J, K));
      D3(D, ILAST, J, K) = D3(D, 1, J, K);
    }
  }
  for (int J = 1; J <= JM; J++) {
    int NR = A1(NROW, J);
    if (A1(KTIPS, NR) <= 0) continue;
    int K1 = A1(KTIPS, NR), K2 = A1(KTIPE, NR);
    if (K1 == 1) K2 = K2 - 1;
    if (K2 == KM) K1 = K1 + 1;
    for (int K = K1; K <= K2; K++) {
      D3(D, 1, J, K) = 0.5f * (D3(D, 1, J, K) + D3(D, ILAST, J, K));
      D3(D, ILAST, J, K) = D3(D, 1, J, K);
    }
  }
  if (o->IFMIX != 0) {
    for (int NR = 1; NR <= NRWSM1; NR++) {
      int J = A1(JMIX, NR), JP1 = J + 1;
      for (int K = 1; K <= KM; K++) {
        float AVGVARJ = 0.0f, AVGVARJP1 = 0.0f;
        for (int I = 1; I <= IMM1; I++) {
          AVGVARJ = AVGVARJ + A1(FP, I) * (D3(D, I, J, K) + D3(D, I + 1, J, K));
          AVGVARJP1 = AVGVARJP1 + A1(FP, I) * (D3(D, I, JP1, K) + D3(D, I + 1, JP1, K));
        }
        AVGVARJ = 0.5f * AVGVARJ;
        AVGVARJP1 = 0.5f * AVGVARJP1;
        for (int I = 1; I <= IM; I++) {
          D3(D, I, J, K) = AVGVARJ;
          D3(D, I, JP1, K) = AVGVARJP1;
        }
      }
    }
  }
}

/* ----------------------------------------------------------------------- */
/* STEPUP, mf:7916-8002 (ideal gas)                                         */

void orc_stepup(orc_t *o, float RELAX) {
  const int IM = o->IM, JM = o->JM, KM = o->KM, IMM1 = IM - 1, KMM1 = KM - 1, KMID = o->KMID;
  float TLIM = 0.1f * A1(TO1, KMID);
  float VLIM = 0.01f * A1(VM1, KMID) * A1(VM1, KMID);
  float RELAX1 = 1.0f - RELAX;
  float *TEMP1 = o->TEMP1;
  for (int K = 1; K <= KM; K++)
    for (int J = 1; J <= JM; J++)
      for (int I = 1; I <= IM; I++) {
        float VSQ = A3(VX, I, J, K) * A3(VX, I, J, K) + A3(VT, I, J, K) * A3(VT, I, J, K) +
                    A3(VR, I, J, K) * A3(VR, I, J, K);
        float HSTAT = A3(HO, I, J, K) - 0.5f * VSQ;
        float TSTATIC = HSTAT * o->RCP;
        float GAMNOW = o->GA;
        if (TSTATIC < TLIM) TSTATIC = TLIM;
        float WTREL = A3(WT, I, J, K);
        float WSQ = VSQ - (A3(VT, I, J, K) * A3(VT, I, J, K) - WTREL * WTREL);
        if (WSQ < VLIM) WSQ = VLIM;
        float V_SONIC = sqrtf(GAMNOW * o->RGAS * TSTATIC);
        float VPLUSC = sqrtf(WSQ) + V_SONIC;
        D3(TEMP1, I, J, K) = o->VSOUND / VPLUSC;
      }
  for (int I = 1; I <= IMM1; I++) {
    int IP1 = I + 1;
    for (int K = 1; K <= KMM1; K++) {
      int KP1 = K + 1;
      float AMACHP = D3(TEMP1, I, 1, K) + D3(TEMP1, IP1, 1, K) + D3(TEMP1, I, 1, KP1) + D3(TEMP1, IP1, 1, KP1);
      for (int J = 2; J <= JM; J++) {
        float AMACHL = D3(TEMP1, I, J, K) + D3(TEMP1, I, J, KP1) + D3(TEMP1, IP1, J, K) + D3(TEMP1, IP1, J, KP1);
        float AVMACH = 0.125f * (AMACHL + AMACHP);
        AMACHP = AMACHL;
        float STEPNEW = A3(BSTEP, I, J, K) * AVMACH;
        A3(STEP, I, J, K) = RELAX * STEPNEW + RELAX1 * A3(STEP, I, J, K);
        A3(RSTEP, I, J, K) = A3(STEP, I, J, K) / A3(BSTEP, I, J, K);
      }
    }
  }
  for (int NR = 1; NR <= o->NROWS - 1; NR++) {
    int J = A1(JMIX, NR), JP1 = J + 1, JP2 = J + 2;
    for (int K = 1; K <= KMM1; K++)
      for (int I = 1; I <= IMM1; I++) {
        A3(STEP, I, JP1, K) = A3(STEP, I, JP2, K);
        A3(RSTEP, I, JP1, K) = 0.0f;
      }
  }
}

/* ----------------------------------------------------------------------- */
/* NEW_MIXPLAN, mf:14922-15251 (ideal gas, ITIMST<5)                        */

void orc_new_mixplan(orc_t *o) {
  const int IM = o->IM, KM = o->KM, IMID = o->IMID;
  float *ALPHAA = (float *)calloc(IM + 2, sizeof(float)), *GAMA = (float *)calloc(IM + 2, sizeof(float));
  float *WREF_JP2 = (float *)calloc(IM + 2, sizeof(float)), *ROREF_JP2 = (float *)calloc(IM + 2, sizeof(float));
  float *PISENT = (float *)calloc(IM + 2, sizeof(float)), *ROISENT = (float *)calloc(IM + 2, sizeof(float));
  float *WMACH_JP2 = (float *)calloc(IM + 2, sizeof(float)), *TISENT = (float *)calloc(IM + 2, sizeof(float));
  float *WIS_JP2 = (float *)calloc(IM + 2, sizeof(float));
  float RFMIX = o->RFMIX, RFMIX1 = 1.0f - RFMIX;
  for (int NR = 1; NR <= o->NROWS - 1; NR++) {
    for (int K = 1; K <= KM; K++) {
      int KSUB = K;
      if (K == KM) KSUB = KM - 1;
      int J = A1(JMIX, NR), JP1, JP2, JP3;
      float F_ANGLE;
      if (A3(FLOWX, IMID, J, KSUB) < 0.0f) {
        JP1 = J + 1; JP2 = J + 2; JP3 = J + 3; F_ANGLE = o->FANGLE;
      } else {
        J = A1(JMIX, NR) + 1;
        JP1 = J - 1; JP2 = J - 2; JP3 = J - 3; F_ANGLE = 0.0f;
      }
      float F_ANGLE1 = 1.0f - F_ANGLE;
      float RJP1 = AJK(R, JP1, K), UJP1 = AJK(UBLADE, JP1, K);
      float ROMIX = A3(RO, IMID, JP1, K);
      float VXMIX = A3(ROVX, IMID, JP1, K) / ROMIX;
      float VRMIX = A3(ROVR, IMID, JP1, K) / ROMIX;
      float VTMIX = A3(RORVT, IMID, JP1, K) / ROMIX / RJP1;
      float ROEMIX = A3(ROE, IMID, JP1, K);
      float VMSQ = VXMIX * VXMIX + VRMIX * VRMIX;
      float VABSQ = VMSQ + VTMIX * VTMIX;
      float WTMIX = VTMIX - UJP1;
      float WMIXSQ = VMSQ + WTMIX * WTMIX;
      float WMIX = sqrtf(WMIXSQ);
      float ALPHAMIX = asinf(WTMIX / WMIX);
      float GAMAMIX = atan2f(VRMIX, VXMIX);
      float TMIX = (ROEMIX / ROMIX - 0.5f * VABSQ) / o->CV;
      float ROTHALPY = o->CP * TMIX + 0.5f * (WMIXSQ - UJP1 * UJP1);
      float GAGAS = o->GA;
      float PMIX = ROMIX * o->RGAS * TMIX;
      int NMACH = 0;
      float UNOWJP2 = 0.0f, RJP2 = 0.0f;
      for (int I = 1; I <= IM; I++) {
        RJP2 = AJK(R, JP2, K);
        UNOWJP2 = AJK(UBLADE, JP2, K);
        float ROJP2 = A3(RO, I, JP2, K);
        float XVELJP2 = A3(ROVX, I, JP2, K) / ROJP2;
        float RVELJP2 = A3(ROVR, I, JP2, K) / ROJP2;
        float TVELJP2 = A3(RORVT, I, JP2, K) / RJP2 / ROJP2;
        float WTJP2 = TVELJP2 - UNOWJP2;
        float VMSQJP2 = XVELJP2 * XVELJP2 + RVELJP2 * RVELJP2;
        float VELSQJP2 = VMSQJP2 + TVELJP2 * TVELJP2;
        float WACT_SQ_JP2 = VMSQJP2 + WTJP2 * WTJP2;
        float WACT_JP2 = sqrtf(WACT_SQ_JP2);
        float RO_ACT_JP2 = ROJP2;
        float T_ACT_JP2 = (A3(ROE, I, JP2, K) / ROJP2 - 0.5f * VELSQJP2) / o->CV;
        float RJP3 = AJK(R, JP3, K), UNOWJP3 = AJK(UBLADE, JP3, K);
        float ROJP3 = A3(RO, I, JP3, K);
        float ROVXJP3 = A3(ROVX, I, JP3, K), ROVRJP3 = A3(ROVR, I, JP3, K);
        float ROVTJP3 = A3(RORVT, I, JP3, K) / RJP3;
        float ROWTJP3 = ROVTJP3 - UNOWJP3 * ROJP3;
        float ROVMSQJP3 = ROVXJP3 * ROVXJP3 + ROVRJP3 * ROVRJP3;
        float ROWRELSQJP3 = ROVMSQJP3 + ROWTJP3 * ROWTJP3;
        float ROWRELJP3 = sqrtf(ROWRELSQJP3);
        float ALPHAJP3 = asinf(ROWTJP3 / ROWRELJP3);
        float GAMAJP3 = atan2f(ROVRJP3, ROVXJP3);
        /* the JMIX+4 angles (mf:15065-15076) are computed but unused */
        float PSTAT = ROJP2 * o->RGAS * T_ACT_JP2;
        float TISENT_JP2 = TMIX * powf(PSTAT / PMIX, o->FGA);
        float WIS_SQ_JP2 = 2 * (ROTHALPY - o->CP * TISENT_JP2 + 0.5f * UNOWJP2 * UNOWJP2);
        float ROISENT_JP2 = PSTAT / o->RGAS / TISENT_JP2;
        ROISENT[I] = ROISENT_JP2;
        PISENT[I] = PSTAT;
        TISENT[I] = TISENT_JP2;
        if (WIS_SQ_JP2 < 1.0f) WIS_SQ_JP2 = 1.0f;
        WIS_JP2[I] = sqrtf(WIS_SQ_JP2);
        WREF_JP2[I] = WACT_JP2;
        ROREF_JP2[I] = RO_ACT_JP2;
        WMACH_JP2[I] = WACT_JP2 / sqrtf(GAGAS * o->RGAS * T_ACT_JP2);
        if (WMACH_JP2[I] > 1.0f) NMACH = NMACH + 1;
        ALPHAA[I] = F_ANGLE1 * ALPHAMIX + F_ANGLE * ALPHAJP3;
        GAMA[I] = F_ANGLE1 * GAMAMIX + F_ANGLE * GAMAJP3;
      }
      if (NMACH != 0) {
        float FROTN = 1.0f;
        if (WTMIX < 0.0f) FROTN = -1.0f;
        for (int I = IMID + 1; I <= IM; I++) {
          int IM1 = I - 1;
          if (WMACH_JP2[I] > 1.001f && WMACH_JP2[IM1] > 1.001f) {
            float DELP = PISENT[I] - PISENT[IM1];
            float DELTAI = sqrtf(WMACH_JP2[I] * WMACH_JP2[I] - 1) / (ROREF_JP2[I] * WREF_JP2[I] * WREF_JP2[I]);
            float DELTAIM1 = sqrtf(WMACH_JP2[IM1] * WMACH_JP2[IM1] - 1) / (ROREF_JP2[IM1] * WREF_JP2[IM1] * WREF_JP2[IM1]);
            ALPHAA[I] = ALPHAA[IM1] + FROTN * DELP * 0.5f * (DELTAI + DELTAIM1);
          }
        }
        for (int I = IMID - 1; I >= 1; I--) {
          int IP1 = I + 1;
          if (WMACH_JP2[I] > 1.001f && WMACH_JP2[IP1] > 1.001f) {
            float DELP = PISENT[I] - PISENT[IP1];
            float DELTAI = sqrtf(WMACH_JP2[I] * WMACH_JP2[I] - 1) / (ROREF_JP2[I] * WREF_JP2[I] * WREF_JP2[I]);
            float DELTAIP1 = sqrtf(WMACH_JP2[IP1] * WMACH_JP2[IP1] - 1) / (ROREF_JP2[IP1] * WREF_JP2[IP1] * WREF_JP2[IP1]);
            ALPHAA[I] = ALPHAA[IP1] + FROTN * DELP * 0.5f * (DELTAI + DELTAIP1);
          }
        }
        float ROTN = ALPHAA[IM] - ALPHAA[1];
        float SUMFP = 0.0f;
        for (int I = 1; I <= IM; I++) {
          ALPHAA[I] = ALPHAA[I] - ROTN * (SUMFP - 0.5f);
          SUMFP = SUMFP + A1(FP, I);
        }
      }
      for (int I = 1; I <= IM; I++) {
        float RO_JP2 = ROISENT[I];
        float WT_JP2 = WIS_JP2[I] * sinf(ALPHAA[I]);
        float VM_JP2 = WIS_JP2[I] * cosf(ALPHAA[I]);
        float VX_JP2 = VM_JP2 * cosf(GAMA[I]);
        float VR_JP2 = VM_JP2 * sinf(GAMA[I]);
        float VT_JP2 = WT_JP2 + UNOWJP2;
        float V_JP2_SQ = VM_JP2 * VM_JP2 + VT_JP2 * VT_JP2;
        float ROE_JP2 = ROISENT[I] * (o->CV * TISENT[I] + 0.5f * V_JP2_SQ);
        A3(RO, I, JP2, K) = RFMIX1 * A3(RO, I, JP2, K) + RFMIX * RO_JP2;
        A3(ROVX, I, JP2, K) = RFMIX1 * A3(ROVX, I, JP2, K) + RFMIX * RO_JP2 * VX_JP2;
        A3(ROVR, I, JP2, K) = RFMIX1 * A3(ROVR, I, JP2, K) + RFMIX * RO_JP2 * VR_JP2;
        A3(RORVT, I, JP2, K) = RFMIX1 * A3(RORVT, I, JP2, K) + RFMIX * RO_JP2 * VT_JP2 * RJP2;
        A3(ROE, I, JP2, K) = RFMIX1 * A3(ROE, I, JP2, K) + RFMIX * ROE_JP2;
      }
    }
  }
  free(ALPHAA); free(GAMA); free(WREF_JP2); free(ROREF_JP2); free(PISENT); free(ROISENT);
  free(WMACH_JP2); free(TISENT); free(WIS_JP2);
}

/* ----------------------------------------------------------------------- */
/* Spalart-Allmaras model: SPAL_LOSS (mf:13586-14827) with GRADVEL and      */
/* GRADCELL (mf:13468-13580).  Restated for KM > 2.  Of the Reynolds-number */
/* diagnostics only REYNOLDS of each row's trailing edge is kept, which the */
/* pressure-gradient source factor reads.                                   */

/* V of a corner-stored field: Gauss-theorem gradient of cell (I,J,K),
 * mf:13468-13510 */
static void gradvel(orc_t *o, const float *V, int I, int J, int K, float *DVDX, float *DVDR, float *DVDT) {
  const int IM = o->IM, JM = o->JM;
  (void)IM; (void)JM;
  int IP1 = I + 1, JM1 = J - 1, KP1 = K + 1;
  float RVOL = 0.25f / A3(VOL, I, J, K);
  float VAVGJ1 = D3(V, I, JM1, K) + D3(V, IP1, JM1, K) + D3(V, IP1, JM1, KP1) + D3(V, I, JM1, KP1);
  float VAVGJ2 = D3(V, I, J, K) + D3(V, IP1, J, K) + D3(V, IP1, J, KP1) + D3(V, I, J, KP1);
  float VAVGI1 = D3(V, I, J, K) + D3(V, I, JM1, K) + D3(V, I, JM1, KP1) + D3(V, I, J, KP1);
  float VAVGI2 = D3(V, IP1, J, K) + D3(V, IP1, JM1, K) + D3(V, IP1, JM1, KP1) + D3(V, IP1, J, KP1);
  float VAVGK1 = D3(V, I, J, K) + D3(V, IP1, J, K) + D3(V, IP1, JM1, K) + D3(V, I, JM1, K);
  float VAVGK2 = D3(V, I, J, KP1) + D3(V, IP1, J, KP1) + D3(V, IP1, JM1, KP1) + D3(V, I, JM1, KP1);
  float VAVGCELL = 0.125f * (VAVGJ1 + VAVGJ2);
  *DVDX = -(VAVGJ2 * A3(AQX, I, J, K) - VAVGJ1 * A3(AQX, I, JM1, K) + VAVGI1 * A3(ABX, I, J, K) -
            VAVGI2 * A3(ABX, IP1, J, K) + VAVGK1 * A3(ASX, I, J, K) - VAVGK2 * A3(ASX, I, J, KP1)) * RVOL;
  *DVDR = -(VAVGJ2 * A3(AQR, I, J, K) - VAVGJ1 * A3(AQR, I, JM1, K) + VAVGI1 * A3(ABR, I, J, K) -
            VAVGI2 * A3(ABR, IP1, J, K) + VAVGK1 * A3(ASR, I, J, K) - VAVGK2 * A3(ASR, I, J, KP1)) * RVOL -
          VAVGCELL / AJK(RAVG_CELL, J, K);
  *DVDT = -(VAVGI1 - VAVGI2) * AJK(ABT, J, K) * RVOL;
}

/* Gradient of a cell-centred field, mf:13516-13580.  Cells beyond J = JM
 * read as zero (the reference reads the unused, zero J = JM+1 layer of its
 * padded array there). */
static float cellv(orc_t *o, const float *V, int I, int J, int K) {
  if (J > o->JM || J < 1) return 0.0f;
  return D3(V, I, J, K);
}

static void gradcell(orc_t *o, const float *V, int I, int J, int K, float *DVDX, float *DVDR, float *DVDT) {
  const int IMM1 = o->IM - 1, KMM1 = o->KM - 1;
  int IP1 = I + 1, JP1 = J + 1, KP1 = K + 1, IM1 = I - 1, JM1 = J - 1, KM1 = K - 1;
  if (I == 1) IM1 = IMM1;
  if (I == IMM1) IP1 = 1;
  if (K == 1) KM1 = 1;
  if (K == KMM1) KP1 = KMM1;
  float RVOL = 0.5f / A3(VOL, I, J, K);
  float v0 = cellv(o, V, I, J, K);
  float VAVGJ1 = v0 + cellv(o, V, I, JM1, K);
  float VAVGJ2 = v0 + cellv(o, V, I, JP1, K);
  float VAVGI1 = v0 + cellv(o, V, IM1, J, K);
  float VAVGI2 = v0 + cellv(o, V, IP1, J, K);
  float VAVGK1 = v0 + cellv(o, V, I, J, KM1);
  float VAVGK2 = v0 + cellv(o, V, I, J, KP1);
  if (A1(IND, J) == 1) {
    if (I == 1) VAVGI1 = 0.0f;
    if (I == IMM1) VAVGI2 = 0.0f;
  }
  if (K == 1 && (o->IBOUND == 0 || o->IBOUND == 2)) VAVGK1 = 0.0f;
  if (K == KMM1 && o->IBOUND <= 1) VAVGK2 = 0.0f;
  *DVDX = -(VAVGJ2 * A3(AQX, I, J, K) - VAVGJ1 * A3(AQX, I, JM1, K) + VAVGI1 * A3(ABX, I, J, K) -
            VAVGI2 * A3(ABX, IP1, J, K) + VAVGK1 * A3(ASX, I, J, K) - VAVGK2 * A3(ASX, I, J, KP1)) * RVOL;
  *DVDR = -(VAVGJ2 * A3(AQR, I, J, K) - VAVGJ1 * A3(AQR, I, JM1, K) + VAVGI1 * A3(ABR, I, J, K) -
            VAVGI2 * A3(ABR, IP1, J, K) + VAVGK1 * A3(ASR, I, J, K) - VAVGK2 * A3(ASR, I, J, KP1)) * RVOL -
          v0 / AJK(RAVG_CELL, J, K);
  *DVDT = -(VAVGI1 - VAVGI2) * AJK(ABT, J, K) * RVOL;
}

/* Shih et al. wall function (NASA/TM-1999-209398), mf:15980-16047: the wall
 * shear stress from the y+ of the previous call, with the streamwise
 * pressure-gradient velocity when YPLUSWALL <= -10.  There |DPDS| < 1 is
 * taken as 1 and written back through DPDS, as the reference's
 * by-reference argument is.  VISLAM is the caller's laminar viscosity.
 * Returns TWALL; *YPLUS_NEW is the y+ the next call starts from. */
static float wallfun(const orc_t *o, float VISLAM, float PERP, float *DPDS, float DENSITY, float YPLUS_OLD,
                     float WREL_WALL, float *YPLUS_NEW) {
  float SGN_DPDS = 0.0f, VPRES = 0.0f, VSTAR = 0.0f, FYPLUSP = 0.0f, FYPLUS = 0.0f, FRACP;
  if (!(o->YPLUSWALL > -10.0f)) {
    if (fabsf(*DPDS) < 1.0f) *DPDS = 1.0f;
    SGN_DPDS = fabsf(*DPDS) / *DPDS;
    VPRES = powf(VISLAM / DENSITY / DENSITY * fabsf(*DPDS), 0.33333f);
    VSTAR = YPLUS_OLD * VISLAM / (DENSITY * PERP);
    if (SGN_DPDS < 0.0f && VPRES > 0.33f * VSTAR) VPRES = 0.33f * VSTAR;
    if (SGN_DPDS > 0.0f && VPRES > 2.0f * VSTAR) VPRES = 2.0f * VSTAR;
    float YPLUS_PRES = VPRES * PERP * DENSITY / VISLAM;
    float YP1 = YPLUS_PRES, YP2 = YP1 * YP1, YP3 = YP2 * YP1, YP4 = YP2 * YP2;
    if (YP1 <= 4.0f) FYPLUSP = 0.5f * YP2 - 7.31E-3f * YP3;
    if (YP1 > 4.0f && YP1 <= 15.0f)
      FYPLUSP = -15.138f + 8.4688f * YP1 - 0.81976f * YP2 + 3.7292E-2f * YP3 - 6.3866E-4f * YP4;
    if (YP1 > 15.0f && YP1 <= 30.0f)
      FYPLUSP = 11.925f + 0.934f * YP1 - 2.7805E-2f * YP2 + 4.6262E-4f * YP3 - 3.1442E-6f * YP4;
    if (YP1 > 30.0f) FYPLUSP = 5.0f * logf(YP1) + 8.0f;
  }
  float YP1 = YPLUS_OLD, YP2 = YP1 * YP1, YP3 = YP2 * YP1, YP4 = YP2 * YP2;
  if (YP1 <= 5.0f) FYPLUS = YP1 + 1.0E-2f * YP2 - 2.9E-03f * YP3;
  if (YP1 > 5.0f && YP1 <= 30.0f) FYPLUS = -0.872f + 1.465f * YP1 - 7.02E-2f * YP2 + 1.66E-03f * YP3 - 1.495E-5f * YP4;
  if (YP1 > 30.0f && YP1 <= 140.0f)
    FYPLUS = 8.6f + 0.1864f * YP1 - 2.006E-3f * YP2 + 1.144E-5f * YP3 - 2.551E-8f * YP4;
  if (YP1 > 140.0f) FYPLUS = 2.439f * logf(YP1) + 5.0f;
  if (o->YPLUSWALL > -10.0f) FRACP = 0.0f;
  else FRACP = SGN_DPDS * (VPRES / VSTAR) * FYPLUSP / FYPLUS;
  float VPLUS = FYPLUS * (1.0f + FRACP);
  VSTAR = WREL_WALL / VPLUS;
  float TWALL = DENSITY * VSTAR * VSTAR;
  *YPLUS_NEW = DENSITY * VSTAR * PERP / VISLAM;
  return TWALL;
}

/* SET_PWALLGRAD, mf:16052-16114: the streamwise pressure gradient of every
 * cell (limited to DPDS_LIM), then smoothed twice along J within each row,
 * with the row's last J planes extrapolated.  The planes beyond JM that the
 * last row's end treatment touches are the reference's unused, zero array
 * padding: they read as zero here and are not stored. */
static void set_pwallgrad(orc_t *o) {
  const int IM = o->IM, JM = o->JM, KM = o->KM, IMM1 = IM - 1, KMM1 = KM - 1, KMID = o->KMID;
  (void)IM;
  float DPDS_LIM = 20.0f * (A1(PO1, KMID) - 0.5f * (o->PDOWN_HUB + o->PDOWN_TIP) / (float)o->NROWS /
                                                (A1(CHORD, 1) + A1(CHORD, o->NROWS)));
  DPDS_LIM = fabsf(DPDS_LIM);
  for (int J = 2; J <= JM; J++)
    for (int K = 1; K <= KMM1; K++)
      for (int I = 1; I <= IMM1; I++) {
        float DPDX, DPDR, DPDT;
        gradvel(o, o->PEFF, I, J, K, &DPDX, &DPDR, &DPDT);
        float VXCELL = A3(VX, I, J, K) + A3(VX, I, J - 1, K) + A3(VX, I + 1, J, K) + A3(VX, I + 1, J - 1, K) +
                       A3(VX, I, J, K + 1) + A3(VX, I, J - 1, K + 1) + A3(VX, I + 1, J, K + 1) + A3(VX, I + 1, J - 1, K + 1);
        float VRCELL = A3(VR, I, J, K) + A3(VR, I, J - 1, K) + A3(VR, I + 1, J, K) + A3(VR, I + 1, J - 1, K) +
                       A3(VR, I, J, K + 1) + A3(VR, I, J - 1, K + 1) + A3(VR, I + 1, J, K + 1) + A3(VR, I + 1, J - 1, K + 1);
        float WTCELL = A3(WT, I, J, K) + A3(WT, I, J - 1, K) + A3(WT, I + 1, J, K) + A3(WT, I + 1, J - 1, K) +
                       A3(WT, I, J, K + 1) + A3(WT, I, J - 1, K + 1) + A3(WT, I + 1, J, K + 1) + A3(WT, I + 1, J - 1, K + 1);
        float WREL_CELL = sqrtf(VXCELL * VXCELL + VRCELL * VRCELL + WTCELL * WTCELL);
        float DPDS = (DPDX * VXCELL + DPDR * VRCELL + DPDT * WTCELL) / WREL_CELL;
        if (fabsf(DPDS) > DPDS_LIM) DPDS = DPDS_LIM * DPDS / fabsf(DPDS);
        A3(DPDS_CELL, I, J, K) = DPDS;
      }
  float *TEMPP = calloc(JM + 8, sizeof(float));
#define DP(J_) ((J_) <= JM ? A3(DPDS_CELL, I, (J_), K) : 0.0f)
#define DPSET(J_, v) do { float v_ = (v); if ((J_) <= JM) A3(DPDS_CELL, I, (J_), K) = v_; } while (0)
  for (int NR = 1; NR <= o->NROWS; NR++) {
    const int J1 = A1(JSTART, NR), J2 = A1(JMIX, NR);
    for (int K = 1; K <= KMM1; K++)
      for (int I = 1; I <= IMM1; I++) {
        for (int pass = 0; pass < 2; pass++) {
          if (pass == 1) {
            for (int J = J1 + 2; J <= J2 - 1; J++) TEMPP[J] = 0.5f * DP(J) + 0.25f * (DP(J - 1) + DP(J + 1));
            TEMPP[J2 + 1] = 0.0f;
            for (int J = J1 + 3; J <= J2 - 2; J++)
              A3(DPDS_CELL, I, J, K) = 0.5f * TEMPP[J] + 0.25f * (TEMPP[J - 1] + TEMPP[J + 1]);
          }
          DPSET(J2 + 1, 0.0f);
          DPSET(J2 + 2, 2 * DP(J2 + 3) - DP(J2 + 4));
          DPSET(J2, 2 * DP(J2 - 1) - DP(J2 - 2));
        }
      }
  }
#undef DP
#undef DPSET
  free(TEMPP);
}

/* log-law skin friction of SPAL_LOSS's wall cells (sums of four node values,
 * hence the 1/16 and 1/128 factors), mf:14221-14268 */
static float sa_cf(float RE, float DPERP, float ROUGH, float ROAVG4, float WAVG, float VISLAM) {
  float RELOG = 1.f / logf(RE);
  float CF;
  if (ROUGH > 1.0e-7f) {
    if (ROUGH > DPERP) ROUGH = DPERP;
    float REK = RE * ROUGH / DPERP;
    REK = REK - 80.0f;
    if (REK < 0.0f) REK = 0.0f;
    float REKSQ = REK * REK;
    float A1 = -.00178493f + .0000814923f * REK + .000000150445f * REKSQ;
    float A2 = .029072f - .001584f * REK - .00000225194f * REKSQ;
    float A3 = .270313f + .0091409f * REK + .00000451537f * REKSQ + .00000000464767f * REKSQ * REK;
    CF = A1 + A2 * RELOG + A3 * RELOG * RELOG;
  } else {
    CF = -0.00178493f + 0.029072f * RELOG + 0.270313f * RELOG * RELOG;
  }
  float CFLAM = 2.0f / RE;
  CF = fmaxf(CF, CFLAM);
  if (RE < 125.f) CF = CFLAM;
  if (ROUGH > 1.0e-7f) {
    float TW = CF * ROAVG4 * WAVG * WAVG / 128.f;
    float VSTAR = sqrtf(TW / (0.25f * ROAVG4));
    float PLUSK = ROUGH * VSTAR * ROAVG4 / VISLAM / 4.f;
    float RELOG10 = log10f(DPERP / ROUGH);
    float FUNCN = 5.75f * RELOG10 + 8.5f;
    float CF_FULL_ROUGH = 2.0f / (FUNCN * FUNCN);
    if (PLUSK > 45.f && CF > CF_FULL_ROUGH) CF = CF_FULL_ROUGH;
  }
  return CF;
}

/* The part SPAL_LOSS and NEW_LOSS share: stresses and heat flows with the
 * final viscosity, wall friction on the blade, hub and casing surfaces, the
 * I- and K-face sums and the body-force update (SPAL_LOSS mf:14123-14819,
 * NEW_LOSS mf:12787-13455).  TXR..QTT hold the strain rates and temperature
 * gradients of each cell and are overwritten with the stresses and heat
 * flows.  The routines differ in two places, selected by newloss: NEW_LOSS's
 * y+ loop on the I = IM surface stops at IM-1 (mf:13035; SPAL_LOSS's,
 * mf:14404, runs to IM), and its hub/casing loop takes the row of its own J
 * (mf:13169) where SPAL_LOSS's reads the row its blade loop left
 * (mf:14536).  Both enter with J-1 = JM-1 and the row of J = JM, the values
 * their earlier loops leave.  Returns the laminar viscosity of the last cell,
 * the value the routines' VISLAM holds on return. */
static float visc_forces(orc_t *o, float *TXR, float *TXX, float *TXT, float *TRR, float *TRT, float *TTT,
                         float *QXX, float *QRR, float *QTT, int newloss) {
  const int IM = o->IM, JM = o->JM, KM = o->KM, IMM1 = IM - 1, KMM1 = KM - 1, IMID = o->IMID, KMID = o->KMID;
  const size_t N = (size_t)IM * JM * KM;
  float *TRX = calloc(N, 4), *TTX = calloc(N, 4), *TTR = calloc(N, 4);
  float *FORCEX = calloc(N, 4), *FORCER = calloc(N, 4), *FORCET = calloc(N, 4), *ESOURCE = calloc(N, 4);
  const float RF_VIS = o->RF_VIS, RF_VIS1 = 1.0f - RF_VIS;
  float VISLAM = o->VISLAM;
  int NRW = A1(NROW, JM), JM1 = JM - 1, KP1 = 0, IP1 = 0;
  (void)KMID;
  /* stresses and heat flows with the final viscosity, mf:14123-14147 */
  for (int K = 1; K <= KMM1; K++)
    for (int J = 2; J <= JM; J++)
      for (int I = 1; I <= IMM1; I++) {
        VISLAM = A3(VISC_LAM, I, J, K);
        float VISTOT = A3(VISC_RAT, I, J, K) * VISLAM;
        float THCOND = o->CP * VISTOT / o->PRANDTL;
        float VISTOT2 = 2.0f * VISTOT;
        D3(TXR, I, J, K) = VISTOT * D3(TXR, I, J, K);
        D3(TXX, I, J, K) = VISTOT2 * D3(TXX, I, J, K);
        D3(TXT, I, J, K) = VISTOT * D3(TXT, I, J, K);
        D3(TRX, I, J, K) = D3(TXR, I, J, K);
        D3(TRR, I, J, K) = VISTOT2 * D3(TRR, I, J, K);
        D3(TRT, I, J, K) = VISTOT * D3(TRT, I, J, K);
        D3(TTX, I, J, K) = D3(TXT, I, J, K);
        D3(TTT, I, J, K) = VISTOT2 * D3(TTT, I, J, K);
        D3(TTR, I, J, K) = D3(TRT, I, J, K);
        D3(QXX, I, J, K) = -THCOND * D3(QXX, I, J, K);
        D3(QRR, I, J, K) = -THCOND * D3(QRR, I, J, K);
        D3(QTT, I, J, K) = -THCOND * D3(QTT, I, J, K);
      }
  const float VISLAM16 = VISLAM * 16.0f;
  if (o->YPLUSWALL < -10.0f) set_pwallgrad(o); /* mf:12812, 14175 */
  /* blade surfaces and the I = const faces, mf:14159-14500.  NRW and JM1
   * keep their last values where the reference does not reset them. */
  for (int K = 1; K <= KMM1; K++) {
    KP1 = K + 1;
    for (int J = 2; J <= JM; J++) {
      if (A1(IND, J) != 0) {
        JM1 = J - 1;
        NRW = A1(NROW, J);
        if (!((A1(KTIPS, NRW) > 0) && (K >= A1(KTIPS, NRW)) && (K < A1(KTIPE, NRW)))) {
          /* I = 1 surface */
          float VXAVG1 = A3(VX, 2, J, K) + A3(VX, 2, JM1, K) + A3(VX, 2, JM1, KP1) + A3(VX, 2, J, KP1);
          float VRAVG1 = A3(VR, 2, J, K) + A3(VR, 2, JM1, K) + A3(VR, 2, JM1, KP1) + A3(VR, 2, J, KP1);
          float WTAVG1 = A3(WT, 2, J, K) + A3(WT, 2, JM1, K) + A3(WT, 2, JM1, KP1) + A3(WT, 2, J, KP1);
          float ROAVG1 = A3(RO, 2, J, K) + A3(RO, 2, JM1, K) + A3(RO, 2, JM1, KP1) + A3(RO, 2, J, KP1);
          float WAVG = sqrtf(WTAVG1 * WTAVG1 + VXAVG1 * VXAVG1 + VRAVG1 * VRAVG1);
          float bx = A3(ABX, 1, J, K), br = A3(ABR, 1, J, K), bt = AJK(ABT, J, K);
          float AREA1 = sqrtf(bx * bx + br * br + bt * bt);
          float DPERP = A3(VOL, 1, J, K) / AREA1;
          float TWALLI1;
          if (o->YPLUSWALL < -0.001f) { /* Shih wall functions, mf:14209-14219 */
            float YPN;
            TWALLI1 = wallfun(o, VISLAM, DPERP, &A3(DPDS_CELL, 1, J, K), 0.25f * ROAVG1, AJK(YPLUS_I1, J, K),
                              0.25f * WAVG, &YPN);
            AJK(YPLUS_I1, J, K) = fminf(1000.0f, YPN);
          } else {
            float RE = DPERP * ROAVG1 * WAVG / VISLAM16;
            float CF = sa_cf(RE, DPERP, A1(ROUGH_L, NRW), ROAVG1, WAVG, VISLAM);
            TWALLI1 = CF * ROAVG1 * WAVG * WAVG / 128.f;
          }
          float VSTAR = sqrtf(TWALLI1 / (0.25f * ROAVG1));
          for (int I = 1; I <= IMID; I++) {
            float YPLS = VSTAR * sqrtf(A3(DWALLSQ, I, J, K)) * (0.25f * ROAVG1) / VISLAM;
            A3(Y_PLUS, I, J, K) = fminf(1000.0f, YPLS);
          }
          float FMULT = -TWALLI1 * AREA1 / WAVG;
          float XFORCE1 = FMULT * VXAVG1, RFORCE1 = FMULT * VRAVG1, TFORCE1 = FMULT * WTAVG1;
          float WVISC1 = TFORCE1 * A1(WRAD, J) * AJK(RAVG_CELL, J, K);
          D3(FORCEX, 1, J, K) = D3(FORCEX, 1, J, K) + XFORCE1;
          D3(FORCER, 1, J, K) = D3(FORCER, 1, J, K) + RFORCE1;
          D3(FORCET, 1, J, K) = D3(FORCET, 1, J, K) + TFORCE1;
          D3(ESOURCE, 1, J, K) = D3(ESOURCE, 1, J, K) + WVISC1;
          /* I = IM surface */
          float VXAVGIM = A3(VX, IMM1, J, K) + A3(VX, IMM1, JM1, K) + A3(VX, IMM1, JM1, KP1) + A3(VX, IMM1, J, KP1);
          float VRAVGIM = A3(VR, IMM1, J, K) + A3(VR, IMM1, JM1, K) + A3(VR, IMM1, JM1, KP1) + A3(VR, IMM1, J, KP1);
          float WTAVGIM = A3(WT, IMM1, J, K) + A3(WT, IMM1, JM1, K) + A3(WT, IMM1, JM1, KP1) + A3(WT, IMM1, J, KP1);
          float ROAVGIM = A3(RO, IMM1, J, K) + A3(RO, IMM1, JM1, K) + A3(RO, IMM1, JM1, KP1) + A3(RO, IMM1, J, KP1);
          WAVG = sqrtf(WTAVGIM * WTAVGIM + VXAVGIM * VXAVGIM + VRAVGIM * VRAVGIM);
          bx = A3(ABX, IM, J, K); br = A3(ABR, IM, J, K);
          float AREAIM = sqrtf(bx * bx + br * br + bt * bt);
          DPERP = A3(VOL, IMM1, J, K) / AREAIM;
          float TWALLIM;
          if (o->YPLUSWALL < -0.001f) { /* mf:14326-14336 */
            float YPN;
            TWALLIM = wallfun(o, VISLAM, DPERP, &A3(DPDS_CELL, IMM1, J, K), 0.25f * ROAVGIM, AJK(YPLUS_IM, J, K),
                              0.25f * WAVG, &YPN);
            AJK(YPLUS_IM, J, K) = fminf(1000.0f, YPN);
          } else {
            float RE = DPERP * ROAVGIM * WAVG / VISLAM16;
            float CF = sa_cf(RE, DPERP, A1(ROUGH_U, NRW), ROAVGIM, WAVG, VISLAM);
            TWALLIM = CF * ROAVGIM * WAVG * WAVG / 128.f;
          }
          VSTAR = sqrtf(TWALLIM / (0.25f * ROAVGIM));
          for (int I = IMID + 1; I <= (newloss ? IMM1 : IM); I++) {
            float YPLS = VSTAR * sqrtf(A3(DWALLSQ, I, J, K)) * (0.25f * ROAVGIM) / VISLAM;
            A3(Y_PLUS, I, J, K) = fminf(1000.f, YPLS);
          }
          FMULT = -TWALLIM * AREAIM / WAVG;
          float XFORCEIM = FMULT * VXAVGIM, RFORCEIM = FMULT * VRAVGIM, TFORCEIM = FMULT * WTAVGIM;
          float WVISCIM = TFORCEIM * A1(WRAD, J) * AJK(RAVG_CELL, J, K);
          D3(FORCEX, IMM1, J, K) = D3(FORCEX, IMM1, J, K) + XFORCEIM;
          D3(FORCER, IMM1, J, K) = D3(FORCER, IMM1, J, K) + RFORCEIM;
          D3(FORCET, IMM1, J, K) = D3(FORCET, IMM1, J, K) + TFORCEIM;
          D3(ESOURCE, IMM1, J, K) = D3(ESOURCE, IMM1, J, K) + WVISCIM;
        }
      }
      /* periodic cells I = 1 and IM-1 (skipped on solid blade surfaces),
       * mf:14434-14486.  The test reads (IND(J)=1 .AND. K<KTIPS) .OR.
       * K>KTIPE, as the reference's operator precedence makes it. */
      if (!(((A1(IND, J) == 1) && (K < A1(KTIPS, NRW))) || (K > A1(KTIPE, NRW)))) {
        int I = 1, IM1 = IMM1;
        float XFOR = 0.5f * ((D3(TXR, I, J, K) + D3(TXR, IM1, J, K)) * A3(ABR, I, J, K) +
                             (D3(TXT, I, J, K) + D3(TXT, IM1, J, K)) * AJK(ABT, J, K) +
                             (D3(TXX, I, J, K) + D3(TXX, IM1, J, K)) * A3(ABX, I, J, K));
        D3(FORCEX, IM1, J, K) = D3(FORCEX, IM1, J, K) + XFOR;
        D3(FORCEX, I, J, K) = D3(FORCEX, I, J, K) - XFOR;
        float AVG = 0.5f * (D3(FORCEX, I, J, K) + D3(FORCEX, IM1, J, K));
        D3(FORCEX, I, J, K) = AVG;
        D3(FORCEX, IM1, J, K) = AVG;
        float RFOR = 0.5f * ((D3(TRR, I, J, K) + D3(TRR, IM1, J, K)) * A3(ABR, I, J, K) +
                             (D3(TRT, I, J, K) + D3(TRT, IM1, J, K)) * AJK(ABT, J, K) +
                             (D3(TRX, I, J, K) + D3(TRX, IM1, J, K)) * A3(ABX, I, J, K));
        D3(FORCER, IM1, J, K) = D3(FORCER, IM1, J, K) + RFOR;
        D3(FORCER, I, J, K) = D3(FORCER, I, J, K) - RFOR;
        AVG = 0.5f * (D3(FORCER, I, J, K) + D3(FORCER, IM1, J, K));
        D3(FORCER, I, J, K) = AVG;
        D3(FORCER, IM1, J, K) = AVG;
        float TFOR = 0.5f * ((D3(TTR, I, J, K) + D3(TTR, IM1, J, K)) * A3(ABR, I, J, K) +
                             (D3(TTT, I, J, K) + D3(TTT, IM1, J, K)) * AJK(ABT, J, K) +
                             (D3(TTX, I, J, K) + D3(TTX, IM1, J, K)) * A3(ABX, I, J, K));
        D3(FORCET, IM1, J, K) = D3(FORCET, IM1, J, K) + TFOR;
        D3(FORCET, I, J, K) = D3(FORCET, I, J, K) - TFOR;
        AVG = 0.5f * (D3(FORCET, I, J, K) + D3(FORCET, IM1, J, K));
        D3(FORCET, I, J, K) = AVG;
        D3(FORCET, IM1, J, K) = AVG;
        float QFLOW = 0.5f * ((D3(QXX, I, J, K) + D3(QXX, IM1, J, K)) * A3(ABX, I, J, K) +
                              (D3(QRR, I, J, K) + D3(QRR, IM1, J, K)) * A3(ABR, I, J, K) +
                              (D3(QTT, I, J, K) + D3(QTT, IM1, J, K)) * AJK(ABT, J, K));
        float VXAVG = A3(VX, I, J, K) + A3(VX, I, JM1, K) + A3(VX, I, JM1, KP1) + A3(VX, I, J, KP1);
        float VRAVG = A3(VR, I, J, K) + A3(VR, I, JM1, K) + A3(VR, I, JM1, KP1) + A3(VR, I, J, KP1);
        float VTAVG = A3(VT, I, J, K) + A3(VT, I, JM1, K) + A3(VT, I, JM1, KP1) + A3(VT, I, J, KP1);
        float WVISC = -0.25f * (XFOR * VXAVG + RFOR * VRAVG + TFOR * VTAVG);
        D3(ESOURCE, I, J, K) = D3(ESOURCE, I, J, K) + QFLOW + WVISC;
        D3(ESOURCE, IM1, J, K) = D3(ESOURCE, IM1, J, K) - QFLOW - WVISC;
        AVG = 0.5f * (D3(ESOURCE, I, J, K) + D3(ESOURCE, IM1, J, K));
        D3(ESOURCE, I, J, K) = AVG;
        D3(ESOURCE, IM1, J, K) = AVG;
      }
      for (int I = 2; I <= IMM1; I++) {
        int IM1 = I - 1;
        float XFOR = 0.5f * ((D3(TXR, I, J, K) + D3(TXR, IM1, J, K)) * A3(ABR, I, J, K) +
                             (D3(TXT, I, J, K) + D3(TXT, IM1, J, K)) * AJK(ABT, J, K) +
                             (D3(TXX, I, J, K) + D3(TXX, IM1, J, K)) * A3(ABX, I, J, K));
        D3(FORCEX, IM1, J, K) = D3(FORCEX, IM1, J, K) + XFOR;
        D3(FORCEX, I, J, K) = D3(FORCEX, I, J, K) - XFOR;
        float RFOR = 0.5f * ((D3(TRR, I, J, K) + D3(TRR, IM1, J, K)) * A3(ABR, I, J, K) +
                             (D3(TRT, I, J, K) + D3(TRT, IM1, J, K)) * AJK(ABT, J, K) +
                             (D3(TRX, I, J, K) + D3(TRX, IM1, J, K)) * A3(ABX, I, J, K));
        D3(FORCER, IM1, J, K) = D3(FORCER, IM1, J, K) + RFOR;
        D3(FORCER, I, J, K) = D3(FORCER, I, J, K) - RFOR;
        float TFOR = 0.5f * ((D3(TTR, I, J, K) + D3(TTR, IM1, J, K)) * A3(ABR, I, J, K) +
                             (D3(TTT, I, J, K) + D3(TTT, IM1, J, K)) * AJK(ABT, J, K) +
                             (D3(TTX, I, J, K) + D3(TTX, IM1, J, K)) * A3(ABX, I, J, K));
        D3(FORCET, IM1, J, K) = D3(FORCET, IM1, J, K) + TFOR;
        D3(FORCET, I, J, K) = D3(FORCET, I, J, K) - TFOR;
        float QFLOW = 0.5f * ((D3(QXX, I, J, K) + D3(QXX, IM1, J, K)) * A3(ABX, I, J, K) +
                              (D3(QRR, I, J, K) + D3(QRR, IM1, J, K)) * A3(ABR, I, J, K) +
                              (D3(QTT, I, J, K) + D3(QTT, IM1, J, K)) * AJK(ABT, J, K));
        float VXAVG = A3(VX, I, J, K) + A3(VX, I, JM1, K) + A3(VX, I, JM1, KP1) + A3(VX, I, J, KP1);
        float VRAVG = A3(VR, I, J, K) + A3(VR, I, JM1, K) + A3(VR, I, JM1, KP1) + A3(VR, I, J, KP1);
        float VTAVG = A3(VT, I, J, K) + A3(VT, I, JM1, K) + A3(VT, I, JM1, KP1) + A3(VT, I, J, KP1);
        float WVISC = -0.25f * (XFOR * VXAVG + RFOR * VRAVG + TFOR * VTAVG);
        D3(ESOURCE, I, J, K) = D3(ESOURCE, I, J, K) + QFLOW + WVISC;
        D3(ESOURCE, IM1, J, K) = D3(ESOURCE, IM1, J, K) - QFLOW - WVISC;
      }
    }
  }
  /* hub and casing surfaces and the K = const faces, mf:14505-14805 */
  for (int J = 2; J <= JM; J++) {
    JM1 = J - 1;
    if (newloss) NRW = A1(NROW, J);
    for (int I = 1; I <= IMM1; I++) {
      IP1 = I + 1;
      float VXAVG1 = A3(VX, I, J, 2) + A3(VX, IP1, J, 2) + A3(VX, IP1, JM1, 2) + A3(VX, I, JM1, 2);
      float VRAVG1 = A3(VR, I, J, 2) + A3(VR, IP1, J, 2) + A3(VR, IP1, JM1, 2) + A3(VR, I, JM1, 2);
      float VTAVG1 = A3(VT, I, J, 2) + A3(VT, IP1, J, 2) + A3(VT, IP1, JM1, 2) + A3(VT, I, JM1, 2);
      float ROAVG1 = A3(RO, I, J, 2) + A3(RO, IP1, J, 2) + A3(RO, IP1, JM1, 2) + A3(RO, I, JM1, 2);
      float RAVG1 = AJK(R, J, 2) + AJK(R, J, 2) + AJK(R, JM1, 2) + AJK(R, JM1, 2);
      float WTAVG1 = VTAVG1 - A1(WHUB, J) * RAVG1;
      float WAVG = sqrtf(VXAVG1 * VXAVG1 + VRAVG1 * VRAVG1 + WTAVG1 * WTAVG1);
      float sx = A3(ASX, I, J, 1), sr = A3(ASR, I, J, 1);
      float AREA1 = sqrtf(sx * sx + sr * sr);
      float DPERP = A3(VOL, I, J, 1) / AREA1;
      float TWALLK1;
      if (o->YPLUSWALL < -0.001f) { /* Shih wall functions, mf:14556-14566 */
        float YPN;
        TWALLK1 = wallfun(o, VISLAM, DPERP, &A3(DPDS_CELL, I, J, 1), 0.25f * ROAVG1, AIJ(YPLUS_K1, I, J),
                          0.25f * WAVG, &YPN);
        AIJ(YPLUS_K1, I, J) = fminf(1000.0f, YPN);
      } else {
        float RE = DPERP * ROAVG1 * WAVG / VISLAM16;
        float CF = sa_cf(RE, DPERP, A1(ROUGH_H, NRW), ROAVG1, WAVG, VISLAM);
        TWALLK1 = CF * ROAVG1 * WAVG * WAVG / 128.f;
      }
      if (o->IBOUND == 0 || o->IBOUND == 2) {
        for (int K = 1; K <= KMID; K++) {
          float VSTAR = sqrtf(TWALLK1 / (0.25f * ROAVG1));
          float YPLS = VSTAR * sqrtf(A3(DWALLSQ, I, J, K)) * (0.25f * ROAVG1) / VISLAM;
          A3(Y_PLUS, I, J, K) = fminf(YPLS, A3(Y_PLUS, I, J, K));
        }
      }
      float FMULT = -TWALLK1 * AREA1 / WAVG;
      if (o->IBOUND == 1 || o->IBOUND > 2) FMULT = 0.0f;
      float XFORCE1 = FMULT * VXAVG1, RFORCE1 = FMULT * VRAVG1, TFORCE1 = FMULT * WTAVG1;
      float WVISC1 = TFORCE1 * 0.25f * A1(WHUB, J) * RAVG1;
      D3(FORCEX, I, J, 1) = D3(FORCEX, I, J, 1) + XFORCE1;
      D3(FORCER, I, J, 1) = D3(FORCER, I, J, 1) + RFORCE1;
      D3(FORCET, I, J, 1) = D3(FORCET, I, J, 1) + TFORCE1;
      D3(ESOURCE, I, J, 1) = D3(ESOURCE, I, J, 1) + WVISC1;
      float VXAVGKM = A3(VX, I, J, KMM1) + A3(VX, IP1, J, KMM1) + A3(VX, IP1, JM1, KMM1) + A3(VX, I, JM1, KMM1);
      float VRAVGKM = A3(VR, I, J, KMM1) + A3(VR, IP1, J, KMM1) + A3(VR, IP1, JM1, KMM1) + A3(VR, I, JM1, KMM1);
      float VTAVGKM = A3(VT, I, J, KMM1) + A3(VT, IP1, J, KMM1) + A3(VT, IP1, JM1, KMM1) + A3(VT, I, JM1, KMM1);
      float ROAVGKM = A3(RO, I, J, KMM1) + A3(RO, IP1, J, KMM1) + A3(RO, IP1, JM1, KMM1) + A3(RO, I, JM1, KMM1);
      float RAVGKM = AJK(R, J, KMM1) + AJK(R, J, KMM1) + AJK(R, JM1, KMM1) + AJK(R, JM1, KMM1);
      float WTAVGKM = VTAVGKM - A1(WTIP, J) * RAVGKM;
      WAVG = sqrtf(VXAVGKM * VXAVGKM + VRAVGKM * VRAVGKM + WTAVGKM * WTAVGKM);
      sx = A3(ASX, I, J, KM); sr = A3(ASR, I, J, KM);
      float AREAKM = sqrtf(sx * sx + sr * sr);
      DPERP = A3(VOL, I, J, KMM1) / AREAKM;
      float TWALLKM;
      if (o->YPLUSWALL < -0.001f) { /* mf:14671-14681 */
        float YPN;
        TWALLKM = wallfun(o, VISLAM, DPERP, &A3(DPDS_CELL, I, J, KMM1), 0.25f * ROAVGKM, AIJ(YPLUS_KM, I, J),
                          0.25f * WAVG, &YPN);
        AIJ(YPLUS_KM, I, J) = fminf(1000.0f, YPN);
      } else {
        float RE = DPERP * ROAVGKM * WAVG / VISLAM16;
        float CF = sa_cf(RE, DPERP, A1(ROUGH_T, NRW), ROAVGKM, WAVG, VISLAM);
        TWALLKM = CF * ROAVGKM * WAVG * WAVG / 128.f;
      }
      if (o->IBOUND == 0 || o->IBOUND == 1) {
        for (int K = KMID + 1; K <= KMM1; K++) {
          float VSTAR = sqrtf(TWALLKM / (0.25f * ROAVGKM));
          float YPLS = VSTAR * sqrtf(A3(DWALLSQ, I, J, K)) * (0.25f * ROAVGKM) / VISLAM;
          A3(Y_PLUS, I, J, K) = fminf(YPLS, A3(Y_PLUS, I, J, K));
        }
      }
      FMULT = -TWALLKM * AREAKM / WAVG;
      if (o->IBOUND >= 2) FMULT = 0.0f;
      float XFORCEKM = FMULT * VXAVGKM, RFORCEKM = FMULT * VRAVGKM, TFORCEKM = FMULT * WTAVGKM;
      float WVISCKM = TFORCEKM * 0.25f * A1(WTIP, J) * RAVGKM;
      D3(FORCEX, I, J, KMM1) = D3(FORCEX, I, J, KMM1) + XFORCEKM;
      D3(FORCER, I, J, KMM1) = D3(FORCER, I, J, KMM1) + RFORCEKM;
      D3(FORCET, I, J, KMM1) = D3(FORCET, I, J, KMM1) + TFORCEKM;
      D3(ESOURCE, I, J, KMM1) = D3(ESOURCE, I, J, KMM1) + WVISCKM;
      for (int K = 2; K <= KMM1; K++) {
        int KM1 = K - 1;
        float XFOR = 0.5f * ((D3(TXR, I, J, K) + D3(TXR, I, J, KM1)) * A3(ASR, I, J, K) +
                             (D3(TXX, I, J, K) + D3(TXX, I, J, KM1)) * A3(ASX, I, J, K));
        D3(FORCEX, I, J, KM1) = D3(FORCEX, I, J, KM1) + XFOR;
        D3(FORCEX, I, J, K) = D3(FORCEX, I, J, K) - XFOR;
        float RFOR = 0.5f * ((D3(TRR, I, J, K) + D3(TRR, I, J, KM1)) * A3(ASR, I, J, K) +
                             (D3(TRX, I, J, K) + D3(TRX, I, J, KM1)) * A3(ASX, I, J, K));
        D3(FORCER, I, J, KM1) = D3(FORCER, I, J, KM1) + RFOR;
        D3(FORCER, I, J, K) = D3(FORCER, I, J, K) - RFOR;
        float TFOR = 0.5f * ((D3(TTR, I, J, K) + D3(TTR, I, J, KM1)) * A3(ASR, I, J, K) +
                             (D3(TTX, I, J, K) + D3(TTX, I, J, KM1)) * A3(ASX, I, J, K));
        D3(FORCET, I, J, KM1) = D3(FORCET, I, J, KM1) + TFOR;
        D3(FORCET, I, J, K) = D3(FORCET, I, J, K) - TFOR;
        float QFLOW = 0.5f * ((D3(QXX, I, J, K) + D3(QXX, I, J, KM1)) * A3(ASX, I, J, K) +
                              (D3(QRR, I, J, K) + D3(QRR, I, J, KM1)) * A3(ASR, I, J, K));
        float VXAVG = A3(VX, I, J, K) + A3(VX, IP1, J, K) + A3(VX, IP1, JM1, K) + A3(VX, I, JM1, K);
        float VRAVG = A3(VR, I, J, K) + A3(VR, IP1, J, K) + A3(VR, IP1, JM1, K) + A3(VR, I, JM1, K);
        float VTAVG = A3(VT, I, J, K) + A3(VT, IP1, J, K) + A3(VT, IP1, JM1, K) + A3(VT, I, JM1, K);
        float WVISC = -0.25f * (XFOR * VXAVG + RFOR * VRAVG + TFOR * VTAVG);
        D3(ESOURCE, I, J, K) = D3(ESOURCE, I, J, K) + QFLOW + WVISC;
        D3(ESOURCE, I, J, KM1) = D3(ESOURCE, I, J, KM1) - QFLOW - WVISC;
      }
    }
  }
  /* body forces, mf:14812-14819 */
  for (int K = 1; K <= KMM1; K++)
    for (int J = 2; J <= JM; J++)
      for (int I = 1; I <= IMM1; I++) {
        A3(XFORCE, I, J, K) = RF_VIS1 * A3(XFORCE, I, J, K) - RF_VIS * D3(FORCEX, I, J, K);
        A3(TFORCE, I, J, K) = RF_VIS1 * A3(TFORCE, I, J, K) - RF_VIS * D3(FORCET, I, J, K) * AJK(RAVG_CELL, J, K);
        A3(RFORCE, I, J, K) = RF_VIS1 * A3(RFORCE, I, J, K) - RF_VIS * D3(FORCER, I, J, K);
        A3(QSOURCE, I, J, K) = RF_VIS1 * A3(QSOURCE, I, J, K) - RF_VIS * D3(ESOURCE, I, J, K);
      }
  free(TRX); free(TTX); free(TTR); free(FORCEX); free(FORCER); free(FORCET); free(ESOURCE);
  return VISLAM;
}

void orc_spal_loss(orc_t *o) {
  const int IM = o->IM, JM = o->JM, KM = o->KM, IMM1 = IM - 1, KMM1 = KM - 1, IMID = o->IMID, KMID = o->KMID;
  const size_t N = (size_t)IM * JM * KM;
  float *TXX = calloc(N, 4), *TXR = calloc(N, 4), *TXT = calloc(N, 4), *TRR = calloc(N, 4), *TRT = calloc(N, 4);
  float *TTT = calloc(N, 4), *QXX = calloc(N, 4), *QRR = calloc(N, 4), *QTT = calloc(N, 4);
  const float RF_VIS = o->RF_VIS, RF_VIS1 = 1.0f - RF_VIS;
  const int NDAMP = 100000;
  float VISLAM = o->VISLAM;
  int NRW = 0, JM1 = 0, KP1 = 0, IP1 = 0;
  /* laminar viscosity over the first quarter of the run, mf:13611-13635 */
  if (o->NSTEP == 1 || o->NSTEP < o->NSTEPS_MAX / 4) {
    if (o->REYNO > 100.f) {
      int J1 = A1(JLE, 1), J2 = A1(JTE, 1);
      if (A1(JLE, 1) > JM) J1 = 1;
      if (A1(JTE, 1) > JM) J2 = JM;
      float XCHORD = AJK(SMERID, J2, KMID) - AJK(SMERID, J1, KMID);
      float a = A3(ROVX, IMID, J2, KMID), b = A3(ROVR, IMID, J2, KMID), c = A3(ROWT, IMID, J2, KMID);
      float ROW2 = sqrtf(a * a + b * b + c * c);
      VISLAM = XCHORD * ROW2 / o->REYNO;
    }
    if (o->REYNO > 0.0f && o->REYNO < 99.99f) VISLAM = o->REYNO / 100000.f;
    if (o->REYNO < 0.0f) VISLAM = -o->REYNO * 1.0e-5f;
    float THCOND = o->CP * VISLAM / o->PRANDTL;
    o->FTCOND = THCOND / VISLAM;
  }
  if (o->FAC_PGRAD > 0.001f) set_pwallgrad(o); /* mf:13656-13658 */
  /* The helicity and pressure-gradient factors of the source (mf:13725-13780)
   * are scalars the cell loop overwrites: the source loop applies the last
   * cell's values to every cell (mf:13954-13955). */
  float VORTICITY_FAC = 1.0f, PGRAD_FAC = 1.0f;
  /* turbulent viscosity of each cell, mf:13662-13888 */
  for (int K = 1; K <= KMM1; K++) {
    KP1 = K + 1;
    for (int J = 2; J <= JM; J++) {
      JM1 = J - 1;
      NRW = A1(NROW, J);
      for (int I = 1; I <= IMM1; I++) {
        IP1 = I + 1;
        float WTAVG = 0.125f * (A3(WT, I, J, K) + A3(WT, IP1, J, K) + A3(WT, IP1, J, KP1) + A3(WT, I, J, KP1) +
                                A3(WT, I, JM1, K) + A3(WT, IP1, JM1, K) + A3(WT, IP1, JM1, KP1) + A3(WT, I, JM1, KP1));
        float ROAVG = A3(ROAVG_CELL, I, J, K);
        float TSAVG = 0.125f * (A3(T_STATIC, I, J, K) + A3(T_STATIC, IP1, J, K) + A3(T_STATIC, IP1, J, KP1) +
                                A3(T_STATIC, I, J, KP1) + A3(T_STATIC, I, JM1, K) + A3(T_STATIC, IP1, JM1, K) +
                                A3(T_STATIC, IP1, JM1, KP1) + A3(T_STATIC, I, JM1, KP1));
        float RAVG = AJK(RAVG_CELL, J, K);
        float DVXDX, DVXDR, DVXDT, DVRDX, DVRDR, DVRDT, DWTDX, DWTDR, DWTDT, DTDX, DTDR, DTDT;
        gradvel(o, o->VX, I, J, K, &DVXDX, &DVXDR, &DVXDT);
        gradvel(o, o->VR, I, J, K, &DVRDX, &DVRDR, &DVRDT);
        gradvel(o, o->WT, I, J, K, &DWTDX, &DWTDR, &DWTDT);
        gradvel(o, o->T_STATIC, I, J, K, &DTDX, &DTDR, &DTDT);
        D3(TXR, I, J, K) = DVXDR + DVRDX;
        D3(TXX, I, J, K) = DVXDX;
        D3(TXT, I, J, K) = DWTDX + DVXDT;
        D3(TRR, I, J, K) = DVRDR;
        D3(TRT, I, J, K) = DWTDR + DVRDT - WTAVG / RAVG;
        D3(TTT, I, J, K) = DWTDT;
        D3(QXX, I, J, K) = DTDX;
        D3(QRR, I, J, K) = DTDR;
        D3(QTT, I, J, K) = DTDT;
        float VORTX = DVRDT - DWTDR - WTAVG / RAVG;
        float VORTR = DWTDX - DVXDT;
        float VORTT = DVXDR - DVRDX;
        float VORT = sqrtf(VORTX * VORTX + VORTR * VORTR + VORTT * VORTT);
        if (J == A1(JSTART, NRW)) VORT = 0.0f;
        if (VORT > o->VORT_MAX) VORT = o->VORT_MAX;
        if (A1(INDMIX, J) == 1) VORT = A3(ABS_VORT, I, J - 1, K);
        A3(ABS_VORT, I, J, K) = VORT;
        if (o->FAC_VORT > 0.001f || o->FAC_PGRAD > 0.001f) { /* mf:13725-13780 */
          float WTSUM = A3(WT, I, J, K) + A3(WT, IP1, J, K) + A3(WT, IP1, J, KP1) + A3(WT, I, J, KP1) +
                        A3(WT, I, JM1, K) + A3(WT, IP1, JM1, K) + A3(WT, IP1, JM1, KP1) + A3(WT, I, JM1, KP1);
          float VXSUM = A3(VX, I, J, K) + A3(VX, IP1, J, K) + A3(VX, IP1, J, KP1) + A3(VX, I, J, KP1) +
                        A3(VX, I, JM1, K) + A3(VX, IP1, JM1, K) + A3(VX, IP1, JM1, KP1) + A3(VX, I, JM1, KP1);
          float VRSUM = A3(VR, I, J, K) + A3(VR, IP1, J, K) + A3(VR, IP1, J, KP1) + A3(VR, I, J, KP1) +
                        A3(VR, I, JM1, K) + A3(VR, IP1, JM1, K) + A3(VR, IP1, JM1, KP1) + A3(VR, I, JM1, KP1);
          float WREL = sqrtf(VXSUM * VXSUM + VRSUM * VRSUM + WTSUM * WTSUM);
          if (o->FAC_VORT > 0.001f) {
            float HELFAC = (VXSUM * VORTX + VRSUM * VORTR + WTSUM * VORTT) / (WREL * VORT);
            VORTICITY_FAC = (1.0f + o->FAC_VORT * tanhf(3.0f * HELFAC * HELFAC));
          }
          if (o->FAC_PGRAD > 0.001f) {
            float a = A3(VX, IMID, J, K), b = A3(VR, IMID, J, K), c = A3(WT, IMID, J, K);
            float WLIM = 0.1f * sqrtf(a * a + b * b + c * c);
            if (WREL < WLIM) WREL = WLIM;
            WREL = 0.125f * WREL;
            float PGRAD = A3(DPDS_CELL, I, J, K);
            if (PGRAD < 0.0f) PGRAD = 0.0f;
            float PTERM = PGRAD * VISLAM / (ROAVG * ROAVG * WREL * WREL * WREL);
            PTERM = PTERM * A1(REYNOLDS, NRW);
            PGRAD_FAC = (1.0f + o->FAC_PGRAD * tanhf(PTERM * PTERM));
          }
        }
        float TURBVIS_WALL = ROAVG * A3(XLENGTH, I, J, K) * VORT * o->FMIXUP;
        if (o->REYNO < 0.0001f) VISLAM = (fabsf(o->REYNO) / 100000.0f) * powf(TSAVG / 288.0f, 0.62f);
        if (I == IMID && K == KMID && J == A1(JTE, NRW)) { /* mf:13795-13806 */
          float a = A3(VX, I, J, K), b = A3(VR, I, J, K), c = A3(WT, I, J, K);
          float WABS = sqrtf(a * a + b * b + c * c);
          A1(ROVEXIT, NRW) = A3(RO, I, J, K) * WABS;
          A1(VISCOSY, NRW) = VISLAM;
          A1(REYNOLDS, NRW) = A1(CHORD, NRW) * A1(ROVEXIT, NRW) / VISLAM;
        }
        float VISTOT = VISLAM * (1 + A3(FST_RAT, I, J, K)) + TURBVIS_WALL;
        float VISLIM = o->TURBVIS_LIM * VISLAM;
        if (VISTOT > VISLIM) VISTOT = VISLIM;
        float VISKLAM = VISLAM / ROAVG;
        A3(DYN_VIS_ML, I, J, K) = (VISTOT - VISLAM);
        if (o->IF_RESTART == 0 && o->NSTEP == 1) {
          float TRANSVISIN;
          if (A1(FSTURB, 1) < 1.0f) TRANSVISIN = 4.35f * VISLAM * powf(A1(FSTURB, 1), 0.25f);
          else TRANSVISIN = VISLAM * (3.5f + A1(FSTURB, 1) * 0.85f);
          A3(TRANS_DYN_VIS, I, J, K) = TRANSVISIN;
        }
        if (o->NSTEP == 1) A3(TRANS_KVIS, I, J, K) = A3(TRANS_DYN_VIS, I, J, K) / ROAVG;
        float TRANSKVIS = A3(TRANS_KVIS, I, J, K);
        float XI = TRANSKVIS / VISKLAM;
        if (XI < 1.1f) XI = 1.1f;
        if (XI > o->TURBVIS_LIM) XI = o->TURBVIS_LIM;
        float XI3 = XI * XI * XI;
        float FV1 = XI3 / (XI3 + 357.9f);
        float TRANSVISLIM = VISLIM / ROAVG / FV1;
        if (TRANSKVIS > TRANSVISLIM) TRANSKVIS = TRANSVISLIM;
        float VISTRU = TRANSKVIS * FV1 * ROAVG;
        A3(TRANS_VORT, I, J, K) = A3(ABS_VORT, I, J, K);
        VISTOT = VISTRU + VISLAM;
        VISLIM = o->TURBVIS_LIM * VISLAM;
        if (VISTOT > VISLIM) VISTOT = VISLIM;
        A3(VISC_RAT, I, J, K) = VISTOT / VISLAM;
        A3(VISC_LAM, I, J, K) = VISLAM;
        A3(VISC_TURB, I, J, K) = VISTOT - VISLAM;
      }
    }
  }
  /* gradient of the turbulent viscosity, mf:13893-13911 */
  for (int K = 1; K <= KMM1; K++)
    for (int J = 2; J <= JM - 1; J++) {
      NRW = A1(NROW, J);
      for (int I = 1; I <= IMM1; I++) {
        float DX_, DR_, DT_;
        gradcell(o, o->TRANS_KVIS, I, J, K, &DX_, &DR_, &DT_);
        A3(GRAD_TVIS, I, J, K) = DX_ * DX_ + DR_ * DR_ + DT_ * DT_;
        float VISCEFF = A3(TRANS_KVIS, I, J, K) + A3(VISC_LAM, I, J, K);
        A3(TVISGRADX, I, J, K) = VISCEFF * DX_;
        A3(TVISGRADR, I, J, K) = VISCEFF * DR_;
        A3(TVISGRADT, I, J, K) = VISCEFF * DT_;
        if (J == A1(JSTART, NRW) || J == A1(JSTART, NRW) - 1) {
          A3(GRAD_TVIS, I, J, K) = 0.0f;
          A3(TVISGRADX, I, J, K) = 0.0f;
          A3(TVISGRADR, I, J, K) = 0.0f;
          A3(TVISGRADT, I, J, K) = 0.0f;
        }
      }
    }
  /* source terms, mf:13916-13997 */
  {
    const float CW2 = 0.3f, SIGMA = 0.6667f, CB2 = 0.622f, CB1 = 0.1355f, CW1 = 0.80607f + 2.433f;
    const float FST0 = o->FAC_ST0 * CB1, FST1 = o->FAC_ST1 / SIGMA, FST2 = o->FAC_ST2 * CB2 / SIGMA;
    const float FST3 = o->FAC_ST3 * CW1, FSTMIX = o->FAC_STMIX / (float)NDAMP;
    float FYPLUS = 0.0f;
    for (int K = 1; K <= KMM1; K++)
      for (int J = 2; J <= JM; J++)
        for (int I = 1; I <= IMM1; I++) {
          float yp = A3(Y_PLUS, I, J, K);
          if (yp <= o->YPLAM) FYPLUS = 0.0f;
          if (yp > o->YPLAM) {
            float XFAC = (yp - o->YPLAM) / (o->YPTURB - o->YPLAM);
            if (XFAC > 1.0f) XFAC = 1.0f;
            FYPLUS = XFAC * XFAC * (3.0f - 2.0f * XFAC);
          }
          float TVISDDX, TVISDDR, TVISDDT, d1, d2;
          gradcell(o, o->TVISGRADX, I, J, K, &TVISDDX, &d1, &d2);
          gradcell(o, o->TVISGRADR, I, J, K, &d1, &TVISDDR, &d2);
          gradcell(o, o->TVISGRADT, I, J, K, &d1, &d2, &TVISDDT);
          float ROAVG = A3(ROAVG_CELL, I, J, K);
          float TVIS = A3(TRANS_KVIS, I, J, K);
          float ST0 = FST0 * A3(TRANS_VORT, I, J, K) * TVIS;
          if (o->FAC_VORT > 0.001f) ST0 = ST0 * VORTICITY_FAC;
          if (o->FAC_PGRAD > 0.001f) ST0 = ST0 * PGRAD_FAC;
          ST0 = ST0 * FYPLUS;
          float ST1 = FST1 * (TVISDDX + TVISDDR + TVISDDT);
          float ST2 = FST2 * A3(GRAD_TVIS, I, J, K);
          float TAU = 5.9488f * TVIS / A3(TRANS_VORT, I, J, K) / A3(DWALLSQ, I, J, K);
          if (TAU > 1.2f) TAU = 1.2f;
          if (TAU < 0.001f) TAU = 0.001f;
          float G = TAU + CW2 * (TAU * TAU * TAU * TAU * TAU * TAU - TAU);
          float G6CW6 = G * G * G * G * G * G / 64.0f;
          float FW = G * powf((1 + 0.015625f) / (1 + G6CW6), 0.166667f);
          float ST3 = FST3 * FW * TVIS * TVIS / A3(DWALLSQ, I, J, K);
          float ST_MIXL = FSTMIX * (A3(DYN_VIS_ML, I, J, K) - A3(VISC_TURB, I, J, K)) / A3(STEP, I, J, K);
          A3(T_SOURCE, I, J, K) = RF_VIS1 * A3(T_SOURCE, I, J, K) +
                                  RF_VIS * (ROAVG * (ST0 + ST1 + ST2 - ST3) * A3(VOL, I, J, K) + ST_MIXL);
          if (J > 2 && A1(INDMIX, J - 1) == 1) A3(T_SOURCE, I, J, K) = 0.0f;
          if (J > 2) {
            if (A1(INDMIX, J - 2) == 1) A3(T_SOURCE, I, J, K) = 0.0f;
          }
        }
  }
  /* transition: damp the source towards laminar in laminar regions,
   * mf:14004-14116 */
  {
    const int IEND1 = IMID / 2, IENDM = IMM1 - IEND1, KEND1 = KMID / 2, KENDM = KMM1 - KEND1;
    const float FTRANS = o->FTRANS;
    for (int J = 2; J <= JM; J++) {
      NRW = A1(NROW, J);
      const int J1 = A1(JSTART, NRW), JROW = J - J1 + 1;
      const int JTRHUB = A1(JTRAN_K1, NRW), JTRTIP = A1(JTRAN_KM, NRW);
      const int JTRLOW = A1(JTRAN_I1, NRW), JTRUP = A1(JTRAN_IM, NRW), JLEROW = A1(JLE, NRW);
      for (int I = 1; I <= IMM1; I++) {
        int damp = 0;
        if (!(JTRHUB <= 1 && FTRANS < 0.1f)) {
          if (JROW < JTRHUB) damp = 1;
          else {
            float RMAX = 0.0f;
            for (int K = 1; K <= KEND1; K++) {
              float RATVIS = A3(VISC_RAT, I, J, K);
              if (RATVIS < RMAX) continue;
              RMAX = RATVIS;
            }
            damp = !(RMAX > FTRANS);
          }
        }
        if (damp) {
          float TVEDGE = A3(VISC_TURB, I, J, KEND1);
          for (int K = 1; K <= KEND1; K++) {
            float FK = (float)(KEND1 + 1 - K) / (float)KEND1;
            FK = FK * FK * (3 - 2 * FK);
            float TVSPEC = FK * TVEDGE;
            A3(T_SOURCE, I, J, K) =
                A3(T_SOURCE, I, J, K) - (A3(VISC_TURB, I, J, K) - TVSPEC) / A3(STEP, I, J, K) / (float)NDAMP;
          }
        }
        damp = 0;
        if (!(JTRTIP <= 1 && FTRANS < 0.1f)) {
          if (JROW < JTRTIP) damp = 1;
          else {
            float RMAX = 0.0f;
            for (int K = KENDM; K <= KMM1; K++) {
              float RATVIS = A3(VISC_RAT, I, J, K);
              if (RATVIS < RMAX) continue;
              RMAX = RATVIS;
            }
            damp = !(RMAX > FTRANS);
          }
        }
        if (damp) {
          float TVEDGE = A3(VISC_TURB, I, J, KENDM);
          for (int K = KENDM; K <= KMM1; K++) {
            float FK = (float)(K + 1 - KENDM) / (float)(KM - KENDM);
            FK = FK * FK * (3 - 2 * FK);
            float TVSPEC = FK * TVEDGE;
            A3(T_SOURCE, I, J, K) =
                A3(T_SOURCE, I, J, K) - (A3(VISC_TURB, I, J, K) - TVSPEC) / A3(STEP, I, J, K) / (float)NDAMP;
          }
        }
      }
      for (int K = 1; K <= KMM1; K++) {
        int damp = 0;
        if (!(JTRLOW <= 1 && FTRANS < 0.1f) && !(J <= JLEROW)) {
          if (JROW < JTRLOW || J < JLEROW) damp = 1;
          else {
            float RMAX = 0.0f;
            for (int I = 1; I <= IEND1; I++) {
              float RATVIS = A3(VISC_RAT, I, J, K);
              if (RATVIS < RMAX) continue;
              RMAX = RATVIS;
            }
            damp = !(RMAX > FTRANS);
          }
        }
        if (damp) {
          float TVEDGE = A3(VISC_TURB, IEND1, J, K);
          for (int I = 1; I <= IEND1; I++) {
            float FI = (float)(IEND1 + 1 - I) / (float)IEND1;
            FI = FI * FI * (3 - 2 * FI);
            float TVSPEC = FI * TVEDGE;
            A3(T_SOURCE, I, J, K) =
                A3(T_SOURCE, I, J, K) - (A3(VISC_TURB, I, J, K) - TVSPEC) / A3(STEP, I, J, K) / (float)NDAMP;
          }
        }
        damp = 0;
        if (!(JTRUP <= 1 && FTRANS < 0.1f) && !(J <= JLEROW)) {
          if (JROW < JTRUP) damp = 1;
          else {
            float RMAX = 0.0f;
            for (int I = IENDM; I <= IMM1; I++) {
              float RATVIS = A3(VISC_RAT, I, J, K);
              if (RATVIS < RMAX) continue;
              RMAX = RATVIS;
            }
            damp = !(RMAX > FTRANS);
          }
        }
        if (damp) {
          float TVEDGE = A3(VISC_TURB, IENDM, J, K);
          for (int I = IENDM; I <= IMM1; I++) {
            float FI = (float)(I + 1 - IENDM) / (float)(IM - IENDM);
            FI = FI * FI * (3 - 2 * FI);
            float TVSPEC = FI * TVEDGE;
            A3(T_SOURCE, I, J, K) =
                A3(T_SOURCE, I, J, K) - (A3(VISC_TURB, I, J, K) - TVSPEC) / A3(STEP, I, J, K) / (float)NDAMP;
          }
        }
      }
    }
  }
  VISLAM = visc_forces(o, TXR, TXX, TXT, TRR, TRT, TTT, QXX, QRR, QTT, 0);
  o->VISLAM = VISLAM;
  free(TXX); free(TXR); free(TXT); free(TRR); free(TRT); free(TTT); free(QXX); free(QRR); free(QTT);
}

/* ----------------------------------------------------------------------- */
/* NEW_LOSS, mf:12522-13462: mixing-length eddy viscosity from the cell      */
/* vorticity, damped through the buffer layer by y+, laminar regions set by  */
/* the transition test, then the stress, wall-friction and force assembly    */
/* it shares with SPAL_LOSS.  Restated for YPLUSWALL >= -0.001 (its Shih     */
/* wall functions, mf:12849-12862, and SET_PWALLGRAD, mf:12812, are not).    */
/* The Reynolds-number diagnostics (VISCOSY, ROVEXIT, REYNOLDS,              */
/* mf:12645-12660) feed printed output only and are left out.                */

void orc_new_loss(orc_t *o) {
  const int IM = o->IM, JM = o->JM, KM = o->KM, IMM1 = IM - 1, KMM1 = KM - 1, IMID = o->IMID, KMID = o->KMID;
  const size_t N = (size_t)IM * JM * KM;
  float *TXX = calloc(N, 4), *TXR = calloc(N, 4), *TXT = calloc(N, 4), *TRR = calloc(N, 4), *TRT = calloc(N, 4);
  float *TTT = calloc(N, 4), *QXX = calloc(N, 4), *QRR = calloc(N, 4), *QTT = calloc(N, 4);
  float VISLAM = o->VISLAM;
  /* laminar viscosity over the first quarter of the run, mf:12544-12570 */
  if (o->NSTEP == 1 || o->NSTEP < o->NSTEPS_MAX / 4) {
    if (o->REYNO > 100.f) {
      int J1 = A1(JLE, 1), J2 = A1(JTE, 1);
      if (A1(JLE, 1) > JM) J1 = 1;
      if (A1(JTE, 1) > JM) J2 = JM;
      float XCHORD = AJK(SMERID, J2, KMID) - AJK(SMERID, J1, KMID);
      float a = A3(ROVX, IMID, J2, KMID), b = A3(ROVR, IMID, J2, KMID), c = A3(ROWT, IMID, J2, KMID);
      float ROW2 = sqrtf(a * a + b * b + c * c);
      VISLAM = XCHORD * ROW2 / o->REYNO;
    }
    if (o->REYNO > 0.0f && o->REYNO < 99.99f) VISLAM = o->REYNO / 100000.f;
    if (o->REYNO < 0.0f) VISLAM = -o->REYNO * 1.0e-5f;
    float THCOND = o->CP * VISLAM / o->PRANDTL;
    o->FTCOND = THCOND / VISLAM;
  }
  /* strain rates and the turbulent viscosity of each cell, mf:12586-12672 */
  float FYPLUS = 0.0f;
  for (int K = 1; K <= KMM1; K++) {
    const int KP1 = K + 1;
    for (int J = 2; J <= JM; J++) {
      const int JM1 = J - 1;
      for (int I = 1; I <= IMM1; I++) {
        const int IP1 = I + 1;
        float WTAVG = 0.125f * (A3(WT, I, J, K) + A3(WT, IP1, J, K) + A3(WT, IP1, J, KP1) + A3(WT, I, J, KP1) +
                                A3(WT, I, JM1, K) + A3(WT, IP1, JM1, K) + A3(WT, IP1, JM1, KP1) + A3(WT, I, JM1, KP1));
        float ROAVG = A3(ROAVG_CELL, I, J, K);
        float TSAVG = 0.125f * (A3(T_STATIC, I, J, K) + A3(T_STATIC, IP1, J, K) + A3(T_STATIC, IP1, J, KP1) +
                                A3(T_STATIC, I, J, KP1) + A3(T_STATIC, I, JM1, K) + A3(T_STATIC, IP1, JM1, K) +
                                A3(T_STATIC, IP1, JM1, KP1) + A3(T_STATIC, I, JM1, KP1));
        float RAVG = AJK(RAVG_CELL, J, K);
        float DVXDX, DVXDR, DVXDT, DVRDX, DVRDR, DVRDT, DWTDX, DWTDR, DWTDT, DTDX, DTDR, DTDT;
        gradvel(o, o->VX, I, J, K, &DVXDX, &DVXDR, &DVXDT);
        gradvel(o, o->VR, I, J, K, &DVRDX, &DVRDR, &DVRDT);
        gradvel(o, o->WT, I, J, K, &DWTDX, &DWTDR, &DWTDT);
        gradvel(o, o->T_STATIC, I, J, K, &DTDX, &DTDR, &DTDT);
        D3(TXR, I, J, K) = DVXDR + DVRDX;
        D3(TXX, I, J, K) = DVXDX;
        D3(TXT, I, J, K) = DWTDX + DVXDT;
        D3(TRR, I, J, K) = DVRDR;
        D3(TRT, I, J, K) = DWTDR + DVRDT - WTAVG / RAVG;
        D3(TTT, I, J, K) = DWTDT;
        D3(QXX, I, J, K) = DTDX;
        D3(QRR, I, J, K) = DTDR;
        D3(QTT, I, J, K) = DTDT;
        float VORTX = DVRDT - DWTDR - WTAVG / RAVG;
        float VORTR = DWTDX - DVXDT;
        float VORTT = DVXDR - DVRDX;
        float VORT = sqrtf(VORTX * VORTX + VORTR * VORTR + VORTT * VORTT);
        if (VORT > o->VORT_MAX) VORT = o->VORT_MAX;
        A3(ABS_VORT, I, J, K) = VORT;
        float TURBVIS_WALL = ROAVG * A3(XLENGTH, I, J, K) * VORT * o->FMIXUP;
        float yp = A3(Y_PLUS, I, J, K);
        if (yp <= o->YPLAM) FYPLUS = 0.0f;
        if (yp > o->YPLAM) {
          float XFAC = (yp - o->YPLAM) / (o->YPTURB - o->YPLAM);
          if (XFAC > 1.0f) XFAC = 1.0f;
          FYPLUS = XFAC * XFAC * (3.0f - 2.0f * XFAC);
        }
        TURBVIS_WALL = TURBVIS_WALL * FYPLUS;
        if (o->REYNO < 0.0001f) VISLAM = (fabsf(o->REYNO) / 100000.0f) * powf(TSAVG / 288.0f, 0.62f);
        float VISTOT = VISLAM * (1 + A3(FST_RAT, I, J, K)) + TURBVIS_WALL;
        float VISLIM = o->TURBVIS_LIM * VISLAM;
        if (VISTOT > VISLIM) VISTOT = VISLIM;
        A3(VISC_RAT, I, J, K) = VISTOT / VISLAM;
        A3(VISC_LAM, I, J, K) = VISLAM;
      }
    }
  }
  /* transition: laminar (viscosity ratio 1) upstream of the transition
   * point of each surface and wherever the largest ratio within half the
   * span or pitch of the wall has not exceeded FTRANS, mf:12677-12772.
   * The hub and casing tests of a column come before the blade-surface
   * tests of the lines, which read what they set. */
  for (int J = 2; J <= JM; J++) {
    const int NRW = A1(NROW, J), JROW = J - A1(JSTART, NRW) + 1;
    const int JTRHUB = A1(JTRAN_K1, NRW), JTRTIP = A1(JTRAN_KM, NRW);
    const int JTRLOW = A1(JTRAN_I1, NRW), JTRUP = A1(JTRAN_IM, NRW);
    if (KM != 2) {
      for (int I = 1; I <= IMM1; I++) {
        int lam = 1;
        if (JROW >= JTRHUB) {
          float RMAX = 0.0f;
          for (int K = 1; K <= KMID; K++) {
            float RATVIS = A3(VISC_RAT, I, J, K);
            if (RATVIS < RMAX) continue;
            RMAX = RATVIS;
          }
          lam = !(RMAX > o->FTRANS);
        }
        if (lam)
          for (int K = 1; K <= KMID; K++) A3(VISC_RAT, I, J, K) = 1.0f;
        lam = 1;
        if (JROW >= JTRTIP) {
          float RMAX = 0.0f;
          for (int K = KMID; K <= KMM1; K++) {
            float RATVIS = A3(VISC_RAT, I, J, K);
            if (RATVIS < RMAX) continue;
            RMAX = RATVIS;
          }
          lam = !(RMAX > o->FTRANS);
        }
        if (lam)
          for (int K = KMID; K <= KMM1; K++) A3(VISC_RAT, I, J, K) = 1.0f;
      }
    }
    for (int K = 1; K <= KMM1; K++) {
      int lam = 1;
      if (JROW >= JTRLOW) {
        float RMAX = 0.0f;
        for (int I = 1; I <= IMID; I++) {
          float RATVIS = A3(VISC_RAT, I, J, K);
          if (RATVIS < RMAX) continue;
          RMAX = RATVIS;
        }
        lam = !(RMAX > o->FTRANS);
      }
      if (lam)
        for (int I = 1; I <= IMID; I++) A3(VISC_RAT, I, J, K) = 1.0f;
      lam = 1;
      if (JROW >= JTRUP) {
        float RMAX = 0.0f;
        for (int I = IMID; I <= IMM1; I++) {
          float RATVIS = A3(VISC_RAT, I, J, K);
          if (RATVIS < RMAX) continue;
          RMAX = RATVIS;
        }
        lam = !(RMAX > o->FTRANS);
      }
      if (lam)
        for (int I = IMID; I <= IMM1; I++) A3(VISC_RAT, I, J, K) = 1.0f;
    }
  }
  o->VISLAM = visc_forces(o, TXR, TXX, TXT, TRR, TRT, TTT, QXX, QRR, QTT, 1);
  free(TXX); free(TXR); free(TXT); free(TRR); free(TRT); free(TTT); free(QXX); free(QRR); free(QTT);
}

/* ----------------------------------------------------------------------- */
/* LOSS, mf:8058-9066: mixing-length eddy viscosity and log-law wall        */
/* friction turned into body forces by a thin-shear-layer approximation.    */
/* Restated for YPLUSWALL >= -0.001 (the Shih wall functions, mf:8269-8280, */
/* and SET_PWALLGRAD are not) and KM > 2.  The entropy-generation rate      */
/* (SGEN/TEMP4, mf:8538-8549 and 9013-9049) and the Reynolds-number and     */
/* VISC_RAT diagnostics feed printed output only and are left out.          */

#define LMAX 1024

/* Skin-friction coefficient of a wall cell, mf:8287-8329: log-law curve fit
 * (with the roughness correction when ROUGH > 1e-7), at least laminar.
 * RO and WB are the density and speed the wall stress is formed with. */
static float loss_cf(float RE, float PERP, float ROUGH, float RO, float WB, float VISLAM) {
  float RELOG = 1.0f / logf(RE);
  float CF;
  if (ROUGH > 1.0e-7f) {
    if (ROUGH > PERP) ROUGH = PERP;
    float REK = RE * ROUGH / PERP;
    REK = REK - 80.0f;
    if (REK < 0.0f) REK = 0.0f;
    float REKSQ = REK * REK;
    float A1 = -.00178493f + .0000814923f * REK + .000000150445f * REKSQ;
    float A2 = .029072f - .001584f * REK - .00000225194f * REKSQ;
    float A3 = .270313f + .0091409f * REK + .00000451537f * REKSQ + .00000000464767f * REKSQ * REK;
    CF = A1 + A2 * RELOG + A3 * RELOG * RELOG;
  } else {
    CF = -0.00178493f + 0.029072f * RELOG + 0.270313f * RELOG * RELOG;
  }
  float CFLAM = 2.0f / RE;
  CF = fmaxf(CF, CFLAM);
  if (RE < 125.f) CF = CFLAM;
  if (ROUGH > 1.0e-7f) {
    float TW = 0.5f * CF * RO * WB * WB;
    float VSTAR = sqrtf(TW / RO);
    float PLUSK = ROUGH * VSTAR * RO / VISLAM;
    float RELOG10 = log10f(PERP / ROUGH);
    float FUNCN = 5.75f * RELOG10 + 8.5f;
    float CF_FULL_ROUGH = 2.0f / (FUNCN * FUNCN);
    if (PLUSK > 45.f && CF > CF_FULL_ROUGH) CF = CF_FULL_ROUGH;
  }
  return CF;
}

/* damping of the turbulent viscosity in the sublayer and buffer region,
 * mf:8496-8501 (FYPLUS keeps its previous value when Y_PLUS is NaN) */
static float loss_fyplus(float yp, float YPLAM, float YPTURB, float prev) {
  float f = prev;
  if (yp <= YPLAM) f = 0.0f;
  if (yp > YPLAM) {
    float XFAC = (yp - YPLAM) / (YPTURB - YPLAM);
    if (XFAC > 1.0f) XFAC = 1.0f;
    f = XFAC * XFAC * (3.0f - 2.0f * XFAC);
  }
  return f;
}

void orc_loss(orc_t *o) {
  const int IM = o->IM, JM = o->JM, KM = o->KM, IMM1 = IM - 1, KMM1 = KM - 1, IMID = o->IMID, KMID = o->KMID;
  float AREA[LMAX], VXAVG[LMAX], VRAVG[LMAX], VTAVG[LMAX], WTAVG[LMAX], ROAVG[LMAX], TAVG[LMAX];
  float RAVG[LMAX], WABS[LMAX], WTB[LMAX], WBOUND[LMAX], XSTRES[LMAX], RSTRES[LMAX], TSTRES[LMAX];
  float QFLOW[LMAX], WVISC[LMAX], VLAM[LMAX], VTURB[LMAX];
  float FYPLUS = 0.0f;
  /* laminar viscosity over the first quarter of the run, mf:8082-8109 */
  if (o->NSTEP == 1 || o->NSTEP < o->NSTEPS_MAX / 4) {
    if (o->REYNO > 100.f) {
      int J1 = A1(JLE, 1), J2 = A1(JTE, 1);
      if (A1(JLE, 1) > JM) J1 = 1;
      if (A1(JTE, 1) > JM) J2 = JM;
      float XCHORD = AJK(SMERID, J2, KMID) - AJK(SMERID, J1, KMID);
      float a = A3(ROVX, IMID, J2, KMID), b = A3(ROVR, IMID, J2, KMID), c = A3(ROWT, IMID, J2, KMID);
      float ROW2 = sqrtf(a * a + b * b + c * c);
      o->VISLAM = XCHORD * ROW2 / o->REYNO;
    }
    if (o->REYNO > 0.0f && o->REYNO < 99.99f) o->VISLAM = o->REYNO / 100000.f;
    if (o->REYNO < 0.0f) o->VISLAM = -o->REYNO * 1.0e-5f;
    o->TCOND = o->CP * o->VISLAM / o->PRANDTL;
    o->FTCOND = o->TCOND / o->VISLAM;
  }
  float VISLAM = o->VISLAM, FTCOND = o->FTCOND;
  const float RF_VIS = o->RF_VIS, RF_VIS1 = 1.f - RF_VIS, FMIXUP = o->FMIXUP;
  /* mixing-length factors, first calls only, mf:8125-8183 */
  if (o->NSTEP <= o->NLOS) {
    for (int NRW = 1; NRW <= o->NROWS; NRW++) {
      float SUMF = 0.0f;
      for (int I = 2; I <= IMM1; I++) {
        SUMF = SUMF + A1(FP, I - 1);
        float XLIM = A1(XLLIM_I1, NRW) * (1.0f - SUMF) + A1(XLLIM_IM, NRW) * SUMF;
        float FPITCH = SUMF * (1.f - SUMF);
        if (FPITCH > XLIM) FPITCH = XLIM;
        float DF = A1(FP, I) + A1(FP, I - 1);
        ANR(FILAM, NRW, I) = A1(FP, I) / DF;
        ANR(FITURB, NRW, I) = 0.16f * FPITCH * FPITCH / (DF * DF);
        ANR(FIWAKE, NRW, I) = 0.16f * A1(XLLIM_DWN, NRW) * A1(XLLIM_DWN, NRW) / (DF * DF);
        ANR(FIUP, NRW, I) = 0.16f * A1(XLLIM_UP, NRW) * A1(XLLIM_UP, NRW) / (DF * DF);
      }
      float DF = A1(FP, 1) + A1(FP, IMM1);
      ANR(FILAM, NRW, 1) = A1(FP, 1) / DF;
      ANR(FIWAKE, NRW, 1) = 0.16f * A1(XLLIM_DWN, NRW) * A1(XLLIM_DWN, NRW) / (DF * DF);
      ANR(FIUP, NRW, 1) = 0.16f * A1(XLLIM_UP, NRW) * A1(XLLIM_UP, NRW) / (DF * DF);
      ANR(FITURB, NRW, 1) = ANR(FITURB, NRW, 2);
      ANR(FILAM, NRW, IM) = A1(FP, IMM1) / DF;
      ANR(FIWAKE, NRW, IM) = 0.16f * A1(XLLIM_DWN, NRW) * A1(XLLIM_DWN, NRW) / (DF * DF);
      ANR(FIUP, NRW, IM) = 0.16f * A1(XLLIM_UP, NRW) * A1(XLLIM_UP, NRW) / (DF * DF);
      ANR(FITURB, NRW, IM) = ANR(FITURB, NRW, IMM1);
      float AVGSPAN = 0.0f;
      const int JMID = (int)(0.5f * (float)(A1(JLE, NRW) + A1(JTE, NRW)));
      for (int K = 2; K <= KM; K++) {
        float XDIF = AJK(X, JMID, K) - AJK(X, JMID, K - 1);
        float RDIF = AJK(R, JMID, K) - AJK(R, JMID, K - 1);
        AVGSPAN = AVGSPAN + sqrtf(XDIF * XDIF + RDIF * RDIF);
      }
      float AVGPIT = 2.0f * 3.1415926f * AJK(R, JMID, KMID) / (float)A1(NBLADE, JMID);
      float XLIMH = A1(XLLIM_K1, NRW) * AVGPIT / AVGSPAN;
      float XLIMT = A1(XLLIM_KM, NRW) * AVGPIT / AVGSPAN;
      SUMF = 0.0f;
      for (int K = 2; K <= KMM1; K++) {
        SUMF = SUMF + A1(FR, K - 1);
        float XLIM = XLIMH * (1.0f - SUMF) + XLIMT * SUMF;
        float FCSPAN = SUMF * (1.f - SUMF);
        if (FCSPAN > XLIM) FCSPAN = XLIM;
        float DF2 = A1(FR, K) + A1(FR, K - 1);
        ANR(FKLAM, NRW, K) = A1(FR, K) / DF2;
        ANR(FKTURB, NRW, K) = 0.16f * FCSPAN * FCSPAN / (DF2 * DF2);
      }
    }
  }
  if (o->YPLUSWALL < -10.0f) set_pwallgrad(o); /* mf:8197 */
  /* stresses on the streamwise (K = const) faces, mf:8201-8556 */
  for (int J = 2; J <= JM; J++) {
    const int NRW = A1(NROW, J), J1 = A1(JSTART, NRW), JROW = J - J1 + 1;
    const int JTRHUB = A1(JTRAN_K1, NRW), JTRTIP = A1(JTRAN_KM, NRW);
    const float WREL = A1(WRAD, J);
    for (int I = 1; I <= IMM1; I++) {
      for (int K = 1; K <= KM; K++) {
        float ax = A3(ASX, I, J, K), ar = A3(ASR, I, J, K);
        AREA[K] = sqrtf(ax * ax + ar * ar);
        VXAVG[K] = 0.25f * (A3(VX, I, J, K) + A3(VX, I, J - 1, K) + A3(VX, I + 1, J, K) + A3(VX, I + 1, J - 1, K));
        VRAVG[K] = 0.25f * (A3(VR, I, J, K) + A3(VR, I, J - 1, K) + A3(VR, I + 1, J, K) + A3(VR, I + 1, J - 1, K));
        VTAVG[K] = 0.25f * (A3(VT, I, J, K) + A3(VT, I, J - 1, K) + A3(VT, I + 1, J, K) + A3(VT, I + 1, J - 1, K));
        ROAVG[K] = 0.25f * (A3(RO, I, J, K) + A3(RO, I, J - 1, K) + A3(RO, I + 1, J, K) + A3(RO, I + 1, J - 1, K));
        TAVG[K] = 0.25f * (A3(T_STATIC, I, J, K) + A3(T_STATIC, I, J - 1, K) + A3(T_STATIC, I + 1, J, K) +
                           A3(T_STATIC, I + 1, J - 1, K));
        RAVG[K] = 0.5f * (AJK(R, J, K) + AJK(R, J - 1, K));
        WTAVG[K] = VTAVG[K] - WREL * RAVG[K];
        float WABSQ = VXAVG[K] * VXAVG[K] + VRAVG[K] * VRAVG[K] + WTAVG[K] * WTAVG[K];
        WABS[K] = sqrtf(WABSQ);
        WTB[K] = VTAVG[K] - A1(WHUB, J) * RAVG[K];
        if (K > KMID) WTB[K] = VTAVG[K] - A1(WTIP, J) * RAVG[K];
        WBOUND[K] = sqrtf(VXAVG[K] * VXAVG[K] + VRAVG[K] * VRAVG[K] + WTB[K] * WTB[K]);
      }
      if (o->REYNO < 0.0001f) {
        VISLAM = (fabsf(o->REYNO) / 100000.0f) * powf(TAVG[KMID] / 288.0f, 0.62f);
        float TCOND = o->CP * VISLAM / o->PRANDTL;
        FTCOND = TCOND / VISLAM;
      }
      float TWALLK1, TWALLKM;
      if (o->YPLUSWALL < -0.001f) { /* Shih wall functions, mf:8269-8280 */
        float YPN;
        TWALLK1 = wallfun(o, VISLAM, A3(VOL, I, J, 1) / AREA[1], &A3(DPDS_CELL, I, J, 1), ROAVG[1], AIJ(YPLUS_K1, I, J),
                          WBOUND[1], &YPN);
        AIJ(YPLUS_K1, I, J) = fminf(1000.0f, YPN);
      } else if (o->YPLUSWALL < 5.f) {
        float PERP = A3(VOL, I, J, 1) / AREA[1];
        float RE = PERP * ROAVG[2] * WBOUND[2] / VISLAM;
        float CF = loss_cf(RE, PERP, A1(ROUGH_H, NRW), ROAVG[2], WBOUND[2], VISLAM);
        TWALLK1 = 0.5f * CF * ROAVG[2] * WBOUND[2] * WBOUND[2];
      } else {
        TWALLK1 = o->CFWALL * ROAVG[1] * WBOUND[1] * WBOUND[1];
      }
      if (o->YPLUSWALL < -0.001f) { /* mf:8348-8358 */
        float YPN;
        TWALLKM = wallfun(o, VISLAM, A3(VOL, I, J, KMM1) / AREA[KM], &A3(DPDS_CELL, I, J, KMM1), ROAVG[KMM1],
                          AIJ(YPLUS_KM, I, J), WBOUND[KMM1], &YPN);
        AIJ(YPLUS_KM, I, J) = fminf(1000.0f, YPN);
      } else if (o->YPLUSWALL < 5.f) {
        float PERP = A3(VOL, I, J, KMM1) / AREA[KM];
        float RE = PERP * ROAVG[KMM1] * WBOUND[KMM1] / VISLAM;
        float CF = loss_cf(RE, PERP, A1(ROUGH_T, NRW), ROAVG[KMM1], WBOUND[KMM1], VISLAM);
        TWALLKM = 0.5f * CF * ROAVG[KMM1] * WBOUND[KMM1] * WBOUND[KMM1];
      } else {
        TWALLKM = o->CFWALL * ROAVG[KM] * WBOUND[KM] * WBOUND[KM];
      }
      /* y+ from the end walls, mf:8420-8435 */
      float VSTAR = sqrtf(TWALLK1 / ROAVG[1]);
      for (int K = 1; K <= KMID; K++) {
        float PERP = sqrtf(A3(DWALLSQ, I, J, K));
        float YPLSK1 = VSTAR * PERP * ROAVG[1] / VISLAM;
        A3(Y_PLUS, I, J, K) = fminf(YPLSK1, 1000.0f);
      }
      VSTAR = sqrtf(TWALLKM / ROAVG[KM]);
      for (int K = KMID + 1; K <= KMM1; K++) {
        float PERP = sqrtf(A3(DWALLSQ, I, J, K));
        float YPLS = VSTAR * PERP * ROAVG[KM] / VISLAM;
        A3(Y_PLUS, I, J, K) = fminf(YPLS, 1000.0f);
      }
      float FMULT = TWALLK1 * AREA[1] / WBOUND[1];
      if (o->IBOUND == 1 || o->IBOUND > 2) FMULT = 0.0f;
      XSTRES[1] = FMULT * VXAVG[1];
      RSTRES[1] = FMULT * VRAVG[1];
      TSTRES[1] = FMULT * WTB[1];
      FMULT = -TWALLKM * AREA[KM] / WBOUND[KM];
      if (o->IBOUND >= 2) FMULT = 0.0f;
      XSTRES[KM] = FMULT * VXAVG[KM];
      RSTRES[KM] = FMULT * VRAVG[KM];
      TSTRES[KM] = FMULT * WTB[KM];
      WVISC[1] = TSTRES[1] * A1(WHUB, J) * RAVG[1];
      QFLOW[1] = 0.0f;
      WVISC[KM] = TSTRES[KM] * A1(WTIP, J) * RAVG[KM];
      QFLOW[KM] = 0.0f;
      for (int K = 2; K <= KMM1; K++) {
        FYPLUS = loss_fyplus(A3(Y_PLUS, I, J, K), o->YPLAM, o->YPTURB, FYPLUS);
        float DPERP = A3(VOL, I, J, K) / AREA[K];
        VLAM[K] = ANR(FKLAM, NRW, K) * VISLAM / DPERP;
        VTURB[K] = ANR(FKTURB, NRW, K) * ROAVG[K] * fabsf(WABS[K + 1] - WABS[K - 1]) * FMIXUP;
        VTURB[K] = VTURB[K] * FYPLUS;
      }
      /* transition, mf:8474-8500 */
      int keep = 0;
      if (JROW >= JTRHUB) {
        float RMAX = 0.0f;
        for (int K = 2; K <= KMID; K++) {
          float RATVIS = VTURB[K] / VLAM[K];
          if (RATVIS < RMAX) continue;
          RMAX = RATVIS;
        }
        keep = RMAX > o->FTRANS;
      }
      if (!keep)
        for (int K = 2; K <= KMID; K++) VTURB[K] = 0.0f;
      keep = 0;
      if (JROW >= JTRTIP) {
        float RMAX = 0.0f;
        for (int K = KMID; K <= KMM1; K++) {
          float RATVIS = VTURB[K] / VLAM[K];
          if (RATVIS < RMAX) continue;
          RMAX = RATVIS;
        }
        keep = RMAX > o->FTRANS;
      }
      if (!keep)
        for (int K = KMID; K <= KMM1; K++) VTURB[K] = 0.0f;
      for (int K = 2; K <= KMM1; K++) {
        float FM = AREA[K] * (VLAM[K] + VTURB[K]);
        XSTRES[K] = FM * (VXAVG[K + 1] - VXAVG[K - 1]);
        RSTRES[K] = FM * (VRAVG[K + 1] - VRAVG[K - 1]);
        float VISTOT = VISLAM * (1.0f + VTURB[K] / VLAM[K]);
        TSTRES[K] = FM * (VTAVG[K + 1] - VTAVG[K - 1]) - VISTOT * AREA[K] * VTAVG[K] / RAVG[K];
        QFLOW[K] = FM * FTCOND * (TAVG[K + 1] - TAVG[K - 1]);
        WVISC[K] = XSTRES[K] * VXAVG[K] + TSTRES[K] * VTAVG[K] + RSTRES[K] * VRAVG[K];
      }
      for (int K = 1; K <= KMM1; K++) {
        A3(XFORCE, I, J, K) = RF_VIS1 * A3(XFORCE, I, J, K) + RF_VIS * (XSTRES[K] - XSTRES[K + 1]);
        A3(RFORCE, I, J, K) = RF_VIS1 * A3(RFORCE, I, J, K) + RF_VIS * (RSTRES[K] - RSTRES[K + 1]);
        A3(TFORCE, I, J, K) =
            RF_VIS1 * A3(TFORCE, I, J, K) + RF_VIS * (TSTRES[K] * RAVG[K] - TSTRES[K + 1] * RAVG[K + 1]);
        A3(QSOURCE, I, J, K) =
            RF_VIS1 * A3(QSOURCE, I, J, K) + RF_VIS * (QFLOW[K] - QFLOW[K + 1] + WVISC[K] - WVISC[K + 1]);
      }
    }
  }
  /* stresses on the blade surfaces and bladewise (I = const) faces,
   * mf:8562-9001 */
  for (int J = 2; J <= JM; J++) {
    const int NRW = A1(NROW, J), J1 = A1(JSTART, NRW), JROW = J - J1 + 1;
    const int JTRLOW = A1(JTRAN_I1, NRW), JTRUP = A1(JTRAN_IM, NRW);
    const int JLEDGE = A1(JLE, NRW), JTEDGE = A1(JTE, NRW);
    const float WREL = A1(WRAD, J);
    for (int K = 1; K <= KMM1; K++) {
      const float RAVGK = AJK(RAVG_CELL, J, K);
      const float abt = AJK(ABT, J, K);
      const bool tipgap = (A1(KTIPS, NRW) > 0) && (K >= A1(KTIPS, NRW)) && (K < A1(KTIPE, NRW));
      for (int I = 1; I <= IM; I++) {
        float bx = A3(ABX, I, J, K), br = A3(ABR, I, J, K);
        AREA[I] = sqrtf(bx * bx + br * br + abt * abt);
        VXAVG[I] = 0.25f * (A3(VX, I, J, K) + A3(VX, I, J - 1, K) + A3(VX, I, J - 1, K + 1) + A3(VX, I, J, K + 1));
        VRAVG[I] = 0.25f * (A3(VR, I, J, K) + A3(VR, I, J - 1, K) + A3(VR, I, J - 1, K + 1) + A3(VR, I, J, K + 1));
        VTAVG[I] = 0.25f * (A3(VT, I, J, K) + A3(VT, I, J - 1, K) + A3(VT, I, J - 1, K + 1) + A3(VT, I, J, K + 1));
        WTAVG[I] = 0.25f * (A3(WT, I, J, K) + A3(WT, I, J - 1, K) + A3(WT, I, J - 1, K + 1) + A3(WT, I, J, K + 1));
        ROAVG[I] = 0.25f * (A3(RO, I, J, K) + A3(RO, I, J - 1, K) + A3(RO, I, J - 1, K + 1) + A3(RO, I, J, K + 1));
        TAVG[I] = 0.25f * (A3(T_STATIC, I, J, K) + A3(T_STATIC, I, J - 1, K) + A3(T_STATIC, I, J - 1, K + 1) +
                           A3(T_STATIC, I, J, K + 1));
        float WABSQ = VXAVG[I] * VXAVG[I] + VRAVG[I] * VRAVG[I] + WTAVG[I] * WTAVG[I];
        WABS[I] = sqrtf(WABSQ);
        WBOUND[I] = WABS[I];
      }
      if (o->REYNO < 0.0001f) {
        VISLAM = (fabsf(o->REYNO) / 100000.0f) * powf(TAVG[IMID] / 288.0f, 0.62f);
        float TCOND = o->CP * VISLAM / o->PRANDTL;
        FTCOND = TCOND / VISLAM;
      }
      if (!(J <= JLEDGE || J > JTEDGE)) {
        /* within the blade row */
        if (!tipgap) {
          float TWALLI1, TWALLIM;
          if (o->YPLUSWALL < -0.001f) { /* Shih wall functions, mf:8621-8631 */
            float YPN;
            TWALLI1 = wallfun(o, VISLAM, A3(VOL, 1, J, K) / AREA[1], &A3(DPDS_CELL, 1, J, K), ROAVG[1],
                              AJK(YPLUS_I1, J, K), WBOUND[2], &YPN);
            AJK(YPLUS_I1, J, K) = fminf(1000.0f, YPN);
          } else if (o->YPLUSWALL < 5.f) {
            float PERP = A3(VOL, 1, J, K) / AREA[1];
            float RE = ROAVG[2] * WABS[2] * PERP / VISLAM;
            float CF = loss_cf(RE, PERP, A1(ROUGH_L, NRW), ROAVG[2], WBOUND[2], VISLAM);
            TWALLI1 = 0.5f * CF * ROAVG[2] * WBOUND[2] * WBOUND[2];
          } else {
            TWALLI1 = o->CFWALL * ROAVG[1] * WABS[1] * WABS[1];
          }
          if (o->YPLUSWALL < -0.001f) { /* mf:8696-8706 */
            float YPN;
            TWALLIM = wallfun(o, VISLAM, A3(VOL, IMM1, J, K) / AREA[IM], &A3(DPDS_CELL, IMM1, J, K), ROAVG[IMM1],
                              AJK(YPLUS_IM, J, K), WBOUND[IMM1], &YPN);
            AJK(YPLUS_IM, J, K) = fminf(1000.0f, YPN);
          } else if (o->YPLUSWALL < 5.f) {
            float PERP = A3(VOL, IMM1, J, K) / AREA[IM];
            float RE = ROAVG[IMM1] * WABS[IMM1] * PERP / VISLAM;
            float CF = loss_cf(RE, PERP, A1(ROUGH_U, NRW), ROAVG[IMM1], WBOUND[IMM1], VISLAM);
            TWALLIM = 0.5f * CF * ROAVG[IMM1] * WBOUND[IMM1] * WBOUND[IMM1];
          } else {
            TWALLIM = o->CFWALL * ROAVG[IM] * WABS[IM] * WABS[IM];
          }
          /* y+ from the blade surfaces, mf:8768-8783 */
          float VSTAR = sqrtf(TWALLI1 / ROAVG[1]);
          for (int I = 1; I <= IMID; I++) {
            float YPLS = VSTAR * sqrtf(A3(DWALLSQ, I, J, K)) * ROAVG[1] / VISLAM;
            float YPLSP = A3(Y_PLUS, I, J, K);
            A3(Y_PLUS, I, J, K) = fminf(YPLS, YPLSP);
          }
          VSTAR = sqrtf(TWALLIM / ROAVG[IM]);
          for (int I = IMID + 1; I <= IMM1; I++) {
            float YPLS = VSTAR * sqrtf(A3(DWALLSQ, I, J, K)) * ROAVG[IM] / VISLAM;
            float YPLSP = A3(Y_PLUS, I, J, K);
            A3(Y_PLUS, I, J, K) = fminf(YPLS, YPLSP);
          }
          float FM = TWALLI1 * AREA[1] / WABS[1];
          XSTRES[1] = FM * VXAVG[1];
          RSTRES[1] = FM * VRAVG[1];
          TSTRES[1] = FM * WTAVG[1];
          FM = -TWALLIM * AREA[IM] / WABS[IM];
          XSTRES[IM] = FM * VXAVG[IM];
          RSTRES[IM] = FM * VRAVG[IM];
          TSTRES[IM] = FM * WTAVG[IM];
          QFLOW[1] = 0.0f;
          WVISC[1] = TSTRES[1] * WREL * RAVGK;
          QFLOW[IM] = 0.0f;
          WVISC[IM] = TSTRES[IM] * WREL * RAVGK;
        }
        for (int I = 2; I <= IMM1; I++) {
          FYPLUS = loss_fyplus(A3(Y_PLUS, I, J, K), o->YPLAM, o->YPTURB, FYPLUS);
          float DPERP = A3(VOL, I, J, K) / AREA[I];
          VLAM[I] = ANR(FILAM, NRW, I) * VISLAM / DPERP;
          VTURB[I] = ANR(FITURB, NRW, I) * ROAVG[I] * fabsf(WABS[I + 1] - WABS[I - 1]) * FMIXUP;
          VTURB[I] = VTURB[I] * FYPLUS;
        }
        int keep = 0;
        if (JROW >= JTRLOW) {
          float RMAX = 0.0f;
          for (int I = 2; I <= IMID; I++) {
            float RATVIS = VTURB[I] / VLAM[I];
            if (RATVIS < RMAX) continue;
            RMAX = RATVIS;
          }
          keep = RMAX > o->FTRANS;
        }
        if (!keep)
          for (int I = 2; I <= IMID; I++) VTURB[I] = 0.0f;
        keep = 0;
        if (JROW >= JTRUP) {
          float RMAX = 0.0f;
          for (int I = IMID; I <= IMM1; I++) {
            float RATVIS = VTURB[I] / VLAM[I];
            if (RATVIS < RMAX) continue;
            RMAX = RATVIS;
          }
          keep = RMAX > o->FTRANS;
        }
        if (!keep)
          for (int I = IMID; I <= IMM1; I++) VTURB[I] = 0.0f;
        for (int I = 2; I <= IMM1; I++) {
          float FM = (VLAM[I] + VTURB[I]) * AREA[I];
          XSTRES[I] = FM * (VXAVG[I + 1] - VXAVG[I - 1]);
          RSTRES[I] = FM * (VRAVG[I + 1] - VRAVG[I - 1]);
          TSTRES[I] = FM * (WTAVG[I + 1] - WTAVG[I - 1]);
          QFLOW[I] = FM * FTCOND * (TAVG[I + 1] - TAVG[I - 1]);
          WVISC[I] = XSTRES[I] * VXAVG[I] + RSTRES[I] * VRAVG[I] + TSTRES[I] * VTAVG[I];
        }
        if (tipgap) {
          XSTRES[1] = 0.5f * (XSTRES[2] + XSTRES[IMM1]);
          RSTRES[1] = 0.5f * (RSTRES[2] + RSTRES[IMM1]);
          TSTRES[1] = 0.5f * (TSTRES[2] + TSTRES[IMM1]);
          XSTRES[IM] = XSTRES[1];
          TSTRES[IM] = TSTRES[1];
          RSTRES[IM] = RSTRES[1];
          QFLOW[1] = 0.5f * (QFLOW[2] + QFLOW[IMM1]);
          WVISC[1] = 0.5f * (WVISC[2] + WVISC[IMM1]);
          QFLOW[IM] = QFLOW[1];
          WVISC[IM] = WVISC[1];
        }
      } else {
        /* upstream of the leading edge and in the wake, mf:8905-8977 */
        const bool wake = J > JTEDGE;
        float FACJ = 0.1f * (float)(J - JTEDGE);
        if (FACJ > 1.f) FACJ = 1.f;
        for (int I = 1; I <= IM; I++) {
          const int ISUB = (I == IM) ? IMM1 : I;
          const int IM1 = (I == 1) ? IMM1 : I - 1, IP1 = (I == IM) ? 2 : I + 1;
          float DPERP = A3(VOL, ISUB, J, K) / AREA[I];
          VLAM[I] = ANR(FILAM, NRW, I) * VISLAM / DPERP;
          if (wake) {
            float FBLEND = (FACJ * ANR(FIWAKE, NRW, I) + (1.f - FACJ) * ANR(FITURB, NRW, I)) * FMIXUP;
            VTURB[I] = FBLEND * ROAVG[I] * fabsf(WABS[IP1] - WABS[IM1]);
          } else {
            VTURB[I] = FMIXUP * ANR(FIUP, NRW, I) * ROAVG[I] * fabsf(WABS[IP1] - WABS[IM1]);
          }
          float FM = AREA[I] * (VLAM[I] + VTURB[I]);
          XSTRES[I] = FM * (VXAVG[IP1] - VXAVG[IM1]);
          RSTRES[I] = FM * (VRAVG[IP1] - VRAVG[IM1]);
          TSTRES[I] = FM * (WTAVG[IP1] - WTAVG[IM1]);
          QFLOW[I] = FM * FTCOND * (TAVG[IP1] - TAVG[IM1]);
          WVISC[I] = XSTRES[I] * VXAVG[I] + TSTRES[I] * VTAVG[I] + RSTRES[I] * VRAVG[I];
        }
        XSTRES[1] = 0.5f * (XSTRES[IM] + XSTRES[1]);
        RSTRES[1] = 0.5f * (RSTRES[IM] + RSTRES[1]);
        TSTRES[1] = 0.5f * (TSTRES[IM] + TSTRES[1]);
        XSTRES[IM] = XSTRES[1];
        RSTRES[IM] = RSTRES[1];
        TSTRES[IM] = TSTRES[1];
        QFLOW[1] = 0.5f * (QFLOW[IM] + QFLOW[1]);
        WVISC[1] = 0.5f * (WVISC[IM] + WVISC[1]);
        QFLOW[IM] = QFLOW[1];
        WVISC[IM] = WVISC[1];
      }
      for (int I = 1; I <= IMM1; I++) {
        A3(XFORCE, I, J, K) = A3(XFORCE, I, J, K) + RF_VIS * (XSTRES[I] - XSTRES[I + 1]);
        A3(RFORCE, I, J, K) = A3(RFORCE, I, J, K) + RF_VIS * (RSTRES[I] - RSTRES[I + 1]);
        A3(TFORCE, I, J, K) = A3(TFORCE, I, J, K) + RF_VIS * (TSTRES[I] - TSTRES[I + 1]) * RAVGK;
        A3(QSOURCE, I, J, K) = A3(QSOURCE, I, J, K) + RF_VIS * (QFLOW[I] - QFLOW[I + 1] + WVISC[I] - WVISC[I + 1]);
      }
    }
  }
  o->VISLAM = VISLAM;
  o->FTCOND = FTCOND;
}

/* ----------------------------------------------------------------------- */
/* one time step: the body of LOOP from label 5000 to the GO TO 5000        */

/* SETFLO, mf:8006-8052: pull the mass flow of every J towards FLOWIN
 * (IN_FLOW = 3) or the mean over J, through the meridional mass flux. */
static void setflo(orc_t *o) {
  const int IM = o->IM, JM = o->JM, KM = o->KM, IMM1 = IM - 1, KMM1 = KM - 1, IMID = o->IMID, KMID = o->KMID;
  float *FLOW = calloc(JM + 1, sizeof(float));
  for (int J = 1; J <= JM; J++) {
    FLOW[J] = 0.0f;
    for (int K = 1; K <= KMM1; K++)
      for (int I = 1; I <= IMM1; I++)
        FLOW[J] = FLOW[J] - A3(FLOWX, I, J, K) * (float)A1(NBLADE, J) / (float)A1(NBLADE, 1);
  }
  float AVFLOW = o->FLOWIN / (float)A1(NBLADE, 1);
  if (o->IN_FLOW != 3) {
    AVFLOW = 0.0f;
    for (int J = 1; J <= JM; J++) AVFLOW = AVFLOW + FLOW[J];
    AVFLOW = AVFLOW / (float)JM;
  }
  for (int J = 1; J <= JM; J++) {
    float ROVMER = sqrtf(A3(ROVX, IMID, J, KMID) * A3(ROVX, IMID, J, KMID) +
                         A3(ROVR, IMID, J, KMID) * A3(ROVR, IMID, J, KMID));
    float DELTA = (AVFLOW / FLOW[J] - 1.0f) * ROVMER;
    DELTA = DELTA * o->RFLOW;
    for (int I = 1; I <= IM; I++)
      for (int K = 1; K <= KM; K++) {
        ROVMER = sqrtf(A3(ROVX, I, J, K) * A3(ROVX, I, J, K) + A3(ROVR, I, J, K) * A3(ROVR, I, J, K));
        if (A3(ROVX, I, J, K) < 0.0f) ROVMER = -ROVMER;
        A3(ROVX, I, J, K) = A3(ROVX, I, J, K) + A3(ROVX, I, J, K) / ROVMER * DELTA;
        A3(ROVR, I, J, K) = A3(ROVR, I, J, K) + A3(ROVR, I, J, K) / ROVMER * DELTA;
        float RO_WT = A3(ROVT, I, J, K) - AJK(UBLADE, J, K) * A3(RO, I, J, K);
        A3(ROVT, I, J, K) = A3(ROVT, I, J, K) + RO_WT / ROVMER * DELTA;
      }
  }
  free(FLOW);
}

static void step_once(orc_t *o) {
  const int IM = o->IM, JM = o->JM, KM = o->KM, IMM1 = IM - 1, KMM1 = KM - 1, JMM1 = JM - 1;
  const int IMID = o->IMID, KMID = o->KMID;
  /* step bookkeeping and smoothing/damping ramp, mf:3753-3824 */
  o->IFPRINT = 0;
  o->NSTEP = o->NSTEP + 1;
  int N = o->NSTEP;
  if (o->NSTEP == o->NSTEPS_MAX) o->IFEND = 1;
  if (o->EAVG < o->CONLIM && o->ECONT < 0.01f) o->IFEND = 1;
  if (fabsf(o->EMAX) < (4 * o->CONLIM) && o->ECONT < 0.01f) o->IFEND = 1;
  if (o->NOUT)
    for (int I = 1; I <= 5; I++)
      if (o->NSTEP == A1(NOUT, I)) o->IFPRINT = 1;
  (void)N;
  float FCHANGE = 1.0f - (float)(o->NSTEP - 1) / (float)o->NCHANGE;
  if (FCHANGE > 0.0f) {
    o->SFT = o->SFTIN + 0.02f * FCHANGE;
    o->SFX = o->SFXIN + 0.02f * FCHANGE;
    o->DAMP = o->DAMPIN * (1.f - 0.75f * FCHANGE);
  }
  float FMIXUP = 1.0f + (o->FACMIXUP - 1.f) * (float)(o->NMIXUP - o->NSTEP) / (float)o->NMIXUP;
  if (FMIXUP < 1.0f) FMIXUP = 1.0f;
  o->FMIXUP = FMIXUP * FMIXUP;

  orc_primitives(o);
  if (o->ILOS == 10 && o->NSTEP % o->NLOS == 0) orc_loss(o);
  if (o->ILOS >= 100 && o->ILOS < 200 && o->NSTEP % o->NLOS == 0) orc_new_loss(o);
  if (o->ILOS >= 200 && o->NSTEP % o->NLOS == 0) orc_spal_loss(o);

  /* mass fluxes, mf:3960-4190 */
  for (int K = 1; K <= KMM1; K++)
    for (int J = 1; J <= JM; J++)
      for (int I = 1; I <= IMM1; I++) {
        float AVGRVX = A3(ROVX, I, J, K) + A3(ROVX, I, J, K + 1) + A3(ROVX, I + 1, J, K) + A3(ROVX, I + 1, J, K + 1);
        float AVGRVR = A3(ROVR, I, J, K) + A3(ROVR, I, J, K + 1) + A3(ROVR, I + 1, J, K) + A3(ROVR, I + 1, J, K + 1);
        A3(FLOWX, I, J, K) = 0.25f * (AVGRVX * A3(AQX, I, J, K) + AVGRVR * A3(AQR, I, J, K));
        A3(SOURCE, I, J, K) = 0.0f;
        A3(XFLUX, I, J, K) = A3(FLOWX, I, J, K);
      }
  for (int K = 1; K <= KM; K++)
    for (int J = 1; J <= JM; J++)
      for (int I = 1; I <= IM; I++) {
        A3(ROWT, I, J, K) = A3(ROVT, I, J, K) - AJK(UBLADE, J, K) * A3(RO, I, J, K);
        A3(WT, I, J, K) = A3(ROWT, I, J, K) / A3(RO, I, J, K);
      }
  for (int K = 1; K <= KMM1; K++)
    for (int J = 2; J <= JM; J++)
      for (int I = 1; I <= IM; I++) {
        float AVGRVX = A3(ROVX, I, J, K) + A3(ROVX, I, J - 1, K) + A3(ROVX, I, J - 1, K + 1) + A3(ROVX, I, J, K + 1);
        float AVGRVT = A3(ROWT, I, J, K) + A3(ROWT, I, J - 1, K) + A3(ROWT, I, J - 1, K + 1) + A3(ROWT, I, J, K + 1);
        float AVGRVR = A3(ROVR, I, J, K) + A3(ROVR, I, J - 1, K) + A3(ROVR, I, J - 1, K + 1) + A3(ROVR, I, J, K + 1);
        A3(FLOWT, I, J, K) = (AVGRVX * A3(ABX, I, J, K) + AVGRVT * AJK(ABT, J, K) + AVGRVR * A3(ABR, I, J, K)) * 0.25f;
        A3(TFLUX, I, J, K) = A3(FLOWT, I, J, K);
      }
  for (int J = 2; J <= JM; J++) {
    if (A1(IND, J) == 1) continue;
    for (int K = 1; K <= KMM1; K++) {
      float AVGFLUX = 0.5f * (A3(FLOWT, 1, J, K) + A3(FLOWT, IM, J, K));
      A3(FLOWT, 1, J, K) = AVGFLUX;
      A3(FLOWT, IM, J, K) = AVGFLUX;
      A3(TFLUX, 1, J, K) = AVGFLUX;
      A3(TFLUX, IM, J, K) = AVGFLUX;
    }
  }
  for (int J = 2; J <= JM; J++) {
    if (A1(IND, J) != 1) continue;
    int NR = A1(NROW, J);
    for (int K = A1(KS1, NR); K <= A1(KS2, NR); K++) {
      A3(TFLUX, 1, J, K) = 0.0f;
      A3(TFLUX, IM, J, K) = 0.0f;
      A3(FLOWT, 1, J, K) = 0.0f;
      A3(FLOWT, IM, J, K) = 0.0f;
    }
  }
  for (int K = 2; K <= KMM1; K++)
    for (int J = 2; J <= JM; J++)
      for (int I = 1; I <= IMM1; I++) {
        float AVGRVX = A3(ROVX, I, J, K) + A3(ROVX, I, J - 1, K) + A3(ROVX, I + 1, J - 1, K) + A3(ROVX, I + 1, J, K);
        float AVGRVR = A3(ROVR, I, J, K) + A3(ROVR, I, J - 1, K) + A3(ROVR, I + 1, J - 1, K) + A3(ROVR, I + 1, J, K);
        A3(FLOWR, I, J, K) = 0.25f * (AVGRVX * A3(ASX, I, J, K) + AVGRVR * A3(ASR, I, J, K));
        A3(RFLUX, I, J, K) = A3(FLOWR, I, J, K);
      }
  for (int J = 2; J <= JM; J++)
    for (int I = 1; I <= IMM1; I++) {
      A3(FLOWR, I, J, 1) = 0.0f;
      A3(FLOWR, I, J, KM) = 0.0f;
      A3(RFLUX, I, J, 1) = 0.0f;
      A3(RFLUX, I, J, KM) = 0.0f;
    }
  orc_tstep(o, o->RO, o->DRO, 1);

  if (o->ILOS >= 100) { /* cell-average density, mf:4260-4270 */
    for (int K = 1; K <= KMM1; K++)
      for (int J = 2; J <= JM; J++)
        for (int I = 1; I <= IMM1; I++) {
          float ROAVG = 0.125f * (A3(RO, I, J, K) + A3(RO, I + 1, J, K) + A3(RO, I + 1, J, K + 1) + A3(RO, I, J, K + 1) +
                                  A3(RO, I, J - 1, K) + A3(RO, I + 1, J - 1, K) + A3(RO, I + 1, J - 1, K + 1) +
                                  A3(RO, I, J - 1, K + 1));
          A3(ROAVG_CELL, I, J, K) = ROAVG;
        }
  }
  if (o->ILOS >= 200) { /* SA transport: doubled fluxes and source, TSTEP NCALL=2, mf:4279-4351 */
    for (int K = 1; K <= KMM1; K++)
      for (int J = 1; J <= JM; J++) {
        int JP1 = J + 1;
        if (J == JM) JP1 = JM;
        if (J == 1) JP1 = 1;
        if (A1(INDMIX, J) == 1) JP1 = J;
        for (int I = 1; I <= IMM1; I++)
          A3(XFLUX, I, J, K) = A3(FLOWX, I, J, K) * (A3(TRANS_KVIS, I, J, K) + A3(TRANS_KVIS, I, JP1, K));
      }
    for (int K = 1; K <= KMM1; K++)
      for (int J = 2; J <= JM; J++)
        for (int I = 1; I <= IMM1; I++) A3(SOURCE, I, J, K) = -2.0f * A3(T_SOURCE, I, J, K);
    for (int K = 1; K <= KMM1; K++)
      for (int J = 2; J <= JM; J++)
        for (int I = 2; I <= IMM1; I++)
          A3(TFLUX, I, J, K) = A3(FLOWT, I, J, K) * (A3(TRANS_KVIS, I - 1, J, K) + A3(TRANS_KVIS, I, J, K));
    for (int K = 1; K <= KMM1; K++)
      for (int J = 2; J <= JM; J++) {
        A3(TFLUX, 1, J, K) = A3(FLOWT, 1, J, K) * (A3(TRANS_KVIS, IMM1, J, K) + A3(TRANS_KVIS, 1, J, K));
        A3(TFLUX, IM, J, K) = A3(TFLUX, 1, J, K);
      }
    for (int K = 2; K <= KMM1; K++)
      for (int J = 2; J <= JM; J++)
        for (int I = 1; I <= IMM1; I++)
          A3(RFLUX, I, J, K) = A3(FLOWR, I, J, K) * (A3(TRANS_KVIS, I, J, K - 1) + A3(TRANS_KVIS, I, J, K));
    for (int J = 2; J <= JM; J++)
      for (int I = 1; I <= IMM1; I++) {
        A3(RFLUX, I, J, 1) = 0.0f;
        A3(RFLUX, I, J, KM) = 0.0f;
      }
    orc_tstep(o, o->TRANS_DYN_VIS, o->DEL_DYNVIS, 2);
    float TRANSVISMIN = 0.1f * A3(VISC_LAM, IMID, 2, KMID);
    for (int K = 1; K <= KMM1; K++)
      for (int J = 2; J <= JM; J++)
        for (int I = 1; I <= IMM1; I++) {
          if (A3(TRANS_DYN_VIS, I, J, K) < TRANSVISMIN) A3(TRANS_DYN_VIS, I, J, K) = TRANSVISMIN;
          A3(TRANS_KVIS, I, J, K) = A3(TRANS_DYN_VIS, I, J, K) / A3(ROAVG_CELL, I, J, K);
        }
  }

  /* energy fluxes, mf:4357-4456 */
  for (int K = 1; K <= KMM1; K++)
    for (int J = 1; J <= JM; J++)
      for (int I = 1; I <= IMM1; I++) {
        float AVGHO = A3(HO, I, J, K) + A3(HO, I, J, K + 1) + A3(HO, I + 1, J, K) + A3(HO, I + 1, J, K + 1);
        A3(XFLUX, I, J, K) = A3(FLOWX, I, J, K) * AVGHO * 0.25f;
      }
  for (int J = 2; J <= JM; J++)
    for (int K = 1; K <= KMM1; K++)
      for (int I = 1; I <= IMM1; I++) A3(SOURCE, I, J, K) = A3(QSOURCE, I, J, K);
  for (int K = 1; K <= KM; K++)
    for (int J = 2; J <= JM; J++)
      for (int I = 1; I <= IMM1; I++) {
        float AVGHO = 0.25f * (A3(HO, I, J, K) + A3(HO, I, J - 1, K) + A3(HO, I + 1, J - 1, K) + A3(HO, I + 1, J, K));
        A3(RFLUX, I, J, K) = A3(FLOWR, I, J, K) * AVGHO;
      }
  for (int K = 1; K <= KMM1; K++)
    for (int J = 2; J <= JM; J++)
      for (int I = 1; I <= IM; I++) {
        float AVGHO = A3(HO, I, J, K) + A3(HO, I, J - 1, K) + A3(HO, I, J - 1, K + 1) + A3(HO, I, J, K + 1);
        float AVGPB = A3(PEFF, I, J, K) + A3(PEFF, I, J - 1, K) + A3(PEFF, I, J - 1, K + 1) + A3(PEFF, I, J, K + 1);
        A3(TFLUX, I, J, K) = (A3(FLOWT, I, J, K) * AVGHO + AJK(WRABT, J, K) * AVGPB) * 0.25f;
      }
  for (int J = 2; J <= JM; J++) {
    if (A1(IND, J) == 1) continue;
    for (int K = 1; K <= KMM1; K++) {
      A3(TFLUX, 1, J, K) = 0.5f * (A3(TFLUX, 1, J, K) + A3(TFLUX, IM, J, K));
      A3(TFLUX, IM, J, K) = A3(TFLUX, 1, J, K);
    }
  }
  orc_tstep(o, o->ROE, o->DROE, 3);

  /* inlet and exit boundary conditions, mf:4476-4757 */
  {
    float PDIFF = 0.0f, SUMP = 0.0f;
    if (o->IPOUT == -1) PDIFF = o->PDOWN_TIP - A1(PD, KM);
    for (int K = 1; K <= KM; K++) {
      float PAVG2 = 0.0f, PAVG3 = 0.0f, PAVGOUT = 0.0f, SUM = 0.0f;
      for (int I = 1; I <= IMM1; I++) {
        PAVG2 = PAVG2 + 0.5f * (A3(P, I, 2, K) + A3(P, I + 1, 2, K)) * A1(FP, I);
        PAVG3 = PAVG3 + 0.5f * (A3(P, I, 3, K) + A3(P, I + 1, 3, K)) * A1(FP, I);
        PAVGOUT = PAVGOUT + 0.5f * (A3(P, I, JMM1, K) + A3(P, I + 1, JMM1, K)) * A1(FP, I);
        SUM = SUM + 0.5f * A1(FP, I) *
                        (A3(ROVT, I, JM, K) * A3(VT, I, JM, K) + A3(ROVT, I + 1, JM, K) * A3(VT, I + 1, JM, K)) /
                        AJK(R, JM, K);
      }
      if (o->IPOUT <= 0) {
        A1(PD, 1) = o->PDOWN_HUB;
        if (K != 1) {
          float DP = (SUM + SUMP) * 0.5f * (AJK(R, JM, K) - AJK(R, JM, K - 1));
          A1(PD, K) = 0.9f * A1(PD, K) + 0.1f * (A1(PD, K - 1) + DP);
        }
        SUMP = SUM;
        if (o->IPOUT == -1) A1(PD, K) = A1(PD, K) + PDIFF;
      }
      for (int I = 1; I <= IM; I++) {
        float RO_STAG = A1(PO1, K) / o->RGAS / A1(TO1, K);
        float RO_IN = A3(RO, I, 1, K);
        if (RO_IN > 0.9999f * RO_STAG) RO_IN = 0.9999f * RO_STAG;
        float PNEW = A1(PO1, K) * powf(RO_IN / RO_STAG, o->GA);
        if (o->IN_PRESS == 0) A3(P, I, 1, K) = o->RFIN1 * A3(P, I, 1, K) + o->RFIN * PNEW;
        if (abs(o->IN_PRESS) == 1)
          A3(P, I, 1, K) = o->RFIN1 * A3(P, I, 1, K) + o->RFIN2 * A3(P, I, 2, K) - o->RFIN * A3(P, I, 3, K);
        if (abs(o->IN_PRESS) == 3) A3(P, I, 1, K) = o->RFIN1 * A3(P, I, 1, K) + o->RFIN2 * PAVG2 - o->RFIN * PAVG3;
        if (abs(o->IN_PRESS) == 4)
          A3(P, I, 1, K) = o->RFIN1 * A3(P, I, 1, K) + o->RFIN2 * A3(P, IMID, 2, KMID) - o->RFIN * A3(P, IMID, 3, KMID);
        if (A3(P, I, 1, K) > A1(PLIM, K)) A3(P, I, 1, K) = 0.9f * A3(P, I, 1, K) + 0.1f * A1(PLIM, K);
        if (A3(P, I, 1, K) > 0.99999f * A1(PO1, K)) A3(P, I, 1, K) = A1(PO1, K) * 0.99999f;
        float TSTATIC = A1(TO1, K) * powf(A3(P, I, 1, K) * A1(RPO1, K), o->FGA);
        float VABSQ = 2 * o->CP * (A1(TO1, K) - TSTATIC);
        A3(T_STATIC, I, 1, K) = TSTATIC;
        float VABS = sqrtf(VABSQ);
        float VMER = 0.0f;
        if (o->IN_VTAN == 0) {
          VMER = VABS * A1(BSCOS, K);
          A3(VT, I, 1, K) = VABS * A1(BSSIN, K);
        }
        if (o->IN_VTAN == 1) {
          float VMERSQ = VABSQ - A1(VTIN, K) * A1(VTIN, K);
          if (VMERSQ < 0.01f) VMERSQ = 0.01f;
          VMER = sqrtf(VMERSQ);
          A3(VT, I, 1, K) = A1(VTIN, K);
        }
        if (o->IN_VTAN == 2) {
          float WR = AJK(UBLADE, 1, K);
          float VRELSQ = VABSQ - WR * WR * A1(BSCOS, K) * A1(BSCOS, K);
          if (VRELSQ < 0.01f) VRELSQ = 0.01f;
          float VREL = sqrtf(VRELSQ) - WR * A1(BSSIN, K);
          VMER = VREL * A1(BSCOS, K);
          A3(VT, I, 1, K) = VREL * A1(BSSIN, K) + WR;
        }
        if (o->IN_VR > 0) A3(VX, I, 1, K) = VMER * A1(BRCOS, K);
        float VXSQ = VMER * VMER - A3(VR, I, 2, K) * A3(VR, I, 2, K);
        if (VXSQ < 0.0000001f * A1(TO1, K) * o->CP) VXSQ = 0.0000001f * A1(TO1, K) * o->CP;
        if (o->IN_VR <= 0) A3(VX, I, 1, K) = sqrtf(VXSQ);
        if (o->IN_VR > 0) A3(VR, I, 1, K) = VMER * A1(BRSIN, K);
        if (o->IN_VR <= 0) A3(VR, I, 1, K) = A3(VR, I, 2, K);
        A3(RO, I, 1, K) = A3(P, I, 1, K) / (o->RGAS * TSTATIC);
        A3(ROVT, I, 1, K) = A3(RO, I, 1, K) * A3(VT, I, 1, K);
        A3(ROVR, I, 1, K) = A3(RO, I, 1, K) * A3(VR, I, 1, K);
        A3(RORVT, I, 1, K) = A3(ROVT, I, 1, K) * AJK(R, 1, K);
        A3(ROVX, I, 1, K) = A3(RO, I, 1, K) * A3(VX, I, 1, K);
        A3(ROE, I, 1, K) = A3(P, I, 1, K) / o->GA1 + A3(RO, I, 1, K) * 0.5f * VABSQ;
        A3(HO, I, 1, K) = o->CP * A1(TO1, K);
        if (o->ILOS >= 200) { /* inlet turbulent viscosity, mf:4679-4690 */
          float TRANSVISIN;
          if (A1(FSTURB, 1) < 1.0f) TRANSVISIN = A3(VISC_LAM, I, 2, K) * 4.35f * powf(A1(FSTURB, 1), 0.25f);
          else TRANSVISIN = A3(VISC_LAM, I, 2, K) * (3.5f + A1(FSTURB, 1) * 0.85f);
          A3(TRANS_DYN_VIS, I, 1, K) = TRANSVISIN;
          A3(TRANS_DYN_VIS, I, 2, K) = A3(TRANS_DYN_VIS, I, 1, K);
          A3(TRANS_KVIS, I, 1, K) = A3(TRANS_DYN_VIS, I, 1, K) / A3(RO, IMID, 1, KMID);
          A3(TRANS_KVIS, I, 2, K) = A3(TRANS_KVIS, I, 1, K);
        }
        /* exit, mf:4690-4730 */
        float EKE = .5f * (A3(VR, I, JM, K) * A3(VR, I, JM, K) + A3(VT, I, JM, K) * A3(VT, I, JM, K) +
                           A3(VX, I, JM, K) * A3(VX, I, JM, K));
        A3(P, I, JM, K) = A1(PD, K) + o->FEXTRAP * (A3(P, I, JMM1, K) - PAVGOUT);
        A3(HO, I, JM, K) = 2.0f * A3(HO, I, JMM1, K) - A3(HO, I, JM - 2, K);
        float TEXIT = (A3(HO, I, JM, K) - EKE) / o->CP;
        A3(RO, I, JM, K) = A3(P, I, JM, K) / o->RGAS / TEXIT;
        A3(ROE, I, JM, K) = A3(RO, I, JM, K) * (o->CV * TEXIT + EKE);
      }
    }
  }
  if (o->IN_FLOW != 0) setflo(o);
  for (int K = 1; K <= KM; K++)
    for (int J = 1; J <= JM; J++) {
      int JP1 = J + 1;
      if (J == JM) JP1 = JM;
      for (int I = 1; I <= IM; I++)
        A3(PEFF, I, J, K) = A3(P, I, J, K) - o->PREFF + o->F_PDOWN * (A3(P, I, JP1, K) - A3(P, I, J, K));
    }

  /* axial momentum, mf:4818-4906 */
  for (int K = 1; K <= KMM1; K++)
    for (int J = 1; J <= JM; J++)
      for (int I = 1; I <= IMM1; I++) {
        A3(SOURCE, I, J, K) = A3(XFORCE, I, J, K);
        float AVGP = A3(PEFF, I, J, K) + A3(PEFF, I, J, K + 1) + A3(PEFF, I + 1, J, K + 1) + A3(PEFF, I + 1, J, K);
        float AVGVX = A3(VX, I, J, K) + A3(VX, I, J, K + 1) + A3(VX, I + 1, J, K + 1) + A3(VX, I + 1, J, K);
        A3(XFLUX, I, J, K) = 0.25f * (AVGP * A3(AQX, I, J, K) + A3(FLOWX, I, J, K) * AVGVX);
      }
  for (int K = 1; K <= KM; K++)
    for (int J = 2; J <= JM; J++)
      for (int I = 1; I <= IMM1; I++) {
        float AVGP = A3(PEFF, I, J, K) + A3(PEFF, I, J - 1, K) + A3(PEFF, I + 1, J - 1, K) + A3(PEFF, I + 1, J, K);
        float AVGVX = A3(VX, I, J, K) + A3(VX, I, J - 1, K) + A3(VX, I + 1, J - 1, K) + A3(VX, I + 1, J, K);
        A3(RFLUX, I, J, K) = 0.25f * (AVGP * A3(ASX, I, J, K) + A3(FLOWR, I, J, K) * AVGVX);
      }
  for (int K = 1; K <= KMM1; K++)
    for (int J = 2; J <= JM; J++)
      for (int I = 1; I <= IM; I++) {
        float AVGVX = A3(VX, I, J, K) + A3(VX, I, J - 1, K) + A3(VX, I, J - 1, K + 1) + A3(VX, I, J, K + 1);
        float AVGP = A3(PEFF, I, J, K) + A3(PEFF, I, J - 1, K) + A3(PEFF, I, J - 1, K + 1) + A3(PEFF, I, J, K + 1);
        A3(TFLUX, I, J, K) = 0.25f * (A3(FLOWT, I, J, K) * AVGVX + A3(ABX, I, J, K) * AVGP);
      }
  for (int J = 2; J <= JM; J++) {
    if (A1(IND, J) == 1) continue;
    for (int K = 1; K <= KMM1; K++) {
      float AVGPCUSP = A3(PEFF, 1, J, K) + A3(PEFF, 1, J - 1, K) + A3(PEFF, 1, J, K + 1) + A3(PEFF, 1, J - 1, K + 1) +
                       A3(PEFF, IM, J, K) + A3(PEFF, IM, J - 1, K) + A3(PEFF, IM, J, K + 1) + A3(PEFF, IM, J - 1, K + 1);
      float AVGVXCUSP = A3(VX, 1, J, K) + A3(VX, 1, J - 1, K) + A3(VX, 1, J, K + 1) + A3(VX, 1, J - 1, K + 1) +
                        A3(VX, IM, J, K) + A3(VX, IM, J - 1, K) + A3(VX, IM, J, K + 1) + A3(VX, IM, J - 1, K + 1);
      A3(TFLUX, 1, J, K) = 0.125f * (A3(FLOWT, 1, J, K) * AVGVXCUSP + AVGPCUSP * A3(ABX, 1, J, K));
      A3(TFLUX, IM, J, K) = 0.125f * (A3(FLOWT, IM, J, K) * AVGVXCUSP + AVGPCUSP * A3(ABX, IM, J, K));
    }
  }
  orc_tstep(o, o->ROVX, o->DROVX, 4);

  /* angular momentum, mf:4913-4996 */
  for (int K = 1; K <= KMM1; K++)
    for (int J = 1; J <= JM; J++) {
      float AVGR = 0.125f * (AJK(R, J, K) + AJK(R, J, K + 1));
      for (int I = 1; I <= IMM1; I++) {
        float AVGVT = A3(VT, I, J, K) + A3(VT, I, J, K + 1) + A3(VT, I + 1, J, K) + A3(VT, I + 1, J, K + 1);
        A3(SOURCE, I, J, K) = A3(TFORCE, I, J, K);
        A3(XFLUX, I, J, K) = A3(FLOWX, I, J, K) * AVGR * AVGVT;
      }
    }
  for (int K = 1; K <= KMM1; K++)
    for (int J = 2; J <= JM; J++) {
      float AVGR = 0.25f * AJK(RAVG_CELL, J, K);
      for (int I = 1; I <= IM; I++) {
        float AVGVT = A3(VT, I, J, K) + A3(VT, I, J - 1, K) + A3(VT, I, J - 1, K + 1) + A3(VT, I, J, K + 1);
        float AVGP = A3(PEFF, I, J, K) + A3(PEFF, I, J - 1, K) + A3(PEFF, I, J - 1, K + 1) + A3(PEFF, I, J, K + 1);
        A3(TFLUX, I, J, K) = (A3(FLOWT, I, J, K) * AVGVT + AJK(ABT, J, K) * AVGP) * AVGR;
      }
    }
  for (int K = 1; K <= KM; K++)
    for (int J = 2; J <= JM; J++) {
      float AVGR = (AJK(R, J, K) + AJK(R, J - 1, K)) * 0.125f;
      for (int I = 1; I <= IMM1; I++) {
        float AVGVT = A3(VT, I, J, K) + A3(VT, I, J - 1, K) + A3(VT, I + 1, J - 1, K) + A3(VT, I + 1, J, K);
        A3(RFLUX, I, J, K) = A3(FLOWR, I, J, K) * AVGVT * AVGR;
      }
    }
  for (int J = 2; J <= JM; J++) {
    if (A1(IND, J) == 1) continue;
    for (int K = 1; K <= KMM1; K++) {
      A3(TFLUX, 1, J, K) = 0.5f * (A3(TFLUX, 1, J, K) + A3(TFLUX, IM, J, K));
      A3(TFLUX, IM, J, K) = A3(TFLUX, 1, J, K);
    }
  }
  orc_tstep(o, o->RORVT, o->DRORVT, 5);

  /* radial momentum, mf:5003-5105 */
  for (int K = 1; K <= KMM1; K++)
    for (int J = 1; J <= JM; J++)
      for (int I = 1; I <= IMM1; I++) {
        float AVGVR = A3(VR, I, J, K) + A3(VR, I, J, K + 1) + A3(VR, I + 1, J, K) + A3(VR, I + 1, J, K + 1);
        float AVGP = A3(PEFF, I, J, K) + A3(PEFF, I, J, K + 1) + A3(PEFF, I + 1, J, K + 1) + A3(PEFF, I + 1, J, K);
        A3(XFLUX, I, J, K) = 0.25f * (A3(FLOWX, I, J, K) * AVGVR + A3(AQR, I, J, K) * AVGP);
      }
  for (int K = 1; K <= KMM1; K++)
    for (int J = 2; J <= JM; J++)
      for (int I = 1; I <= IM; I++) {
        float AVGVR = A3(VR, I, J, K) + A3(VR, I, J - 1, K) + A3(VR, I, J - 1, K + 1) + A3(VR, I, J, K + 1);
        float AVGP = A3(PEFF, I, J, K) + A3(PEFF, I, J - 1, K) + A3(PEFF, I, J - 1, K + 1) + A3(PEFF, I, J, K + 1);
        A3(TFLUX, I, J, K) = .25f * (A3(FLOWT, I, J, K) * AVGVR + AVGP * A3(ABR, I, J, K));
      }
  for (int K = 1; K <= KM; K++)
    for (int J = 2; J <= JM; J++)
      for (int I = 1; I <= IMM1; I++) {
        float AVGVR = A3(VR, I, J, K) + A3(VR, I, J - 1, K) + A3(VR, I + 1, J - 1, K) + A3(VR, I + 1, J, K);
        float AVGP = A3(PEFF, I, J, K) + A3(PEFF, I, J - 1, K) + A3(PEFF, I + 1, J - 1, K) + A3(PEFF, I + 1, J, K);
        A3(RFLUX, I, J, K) = 0.25f * (A3(FLOWR, I, J, K) * AVGVR + A3(ASR, I, J, K) * AVGP);
      }
  for (int K = 1; K <= KMM1; K++)
    for (int J = 2; J <= JM; J++)
      for (int I = 1; I <= IMM1; I++) {
        float AVGRVT = A3(ROVT, I, J, K) + A3(ROVT, I, J, K + 1) + A3(ROVT, I, J - 1, K) + A3(ROVT, I, J - 1, K + 1) +
                       A3(ROVT, I + 1, J, K) + A3(ROVT, I + 1, J, K + 1) + A3(ROVT, I + 1, J - 1, K) +
                       A3(ROVT, I + 1, J - 1, K + 1);
        float AVGVT = 0.125f * (A3(VT, I, J, K) + A3(VT, I, J - 1, K) + A3(VT, I, J - 1, K + 1) + A3(VT, I, J, K + 1) +
                                A3(VT, I + 1, J, K) + A3(VT, I + 1, J - 1, K) + A3(VT, I + 1, J - 1, K + 1) +
                                A3(VT, I + 1, J, K + 1));
        float AVGP = A3(PEFF, I, J, K) + A3(PEFF, I, J - 1, K) + A3(PEFF, I, J - 1, K + 1) + A3(PEFF, I, J, K + 1) +
                     A3(PEFF, I + 1, J, K) + A3(PEFF, I + 1, J - 1, K) + A3(PEFF, I + 1, J - 1, K + 1) +
                     A3(PEFF, I + 1, J, K + 1);
        A3(SOURCE, I, J, K) = A3(RFORCE, I, J, K) - (AVGRVT * AVGVT + AVGP) * A3(VOLOR, I, J, K);
      }
  for (int J = 2; J <= JM; J++) {
    if (A1(IND, J) == 1) continue;
    for (int K = 1; K <= KMM1; K++) {
      float AVGP1 = 0.25f * (A3(PEFF, 1, J, K) + A3(PEFF, 1, J - 1, K) + A3(PEFF, 1, J, K + 1) + A3(PEFF, 1, J - 1, K + 1));
      float AVGPM = 0.25f * (A3(PEFF, IM, J, K) + A3(PEFF, IM, J - 1, K) + A3(PEFF, IM, J, K + 1) + A3(PEFF, IM, J - 1, K + 1));
      float T1 = A3(TFLUX, 1, J, K) - AVGP1 * A3(ABR, 1, J, K);
      float T2 = A3(TFLUX, IM, J, K) - AVGPM * A3(ABR, IM, J, K);
      A3(TFLUX, 1, J, K) = 0.5f * ((T1 + T2) + (AVGP1 + AVGPM) * A3(ABR, 1, J, K));
      A3(TFLUX, IM, J, K) = 0.5f * ((T1 + T2) + (AVGP1 + AVGPM) * A3(ABR, IM, J, K));
    }
  }
  orc_tstep(o, o->ROVR, o->DROVR, 6);

  /* density and Mach limiters, mf:5115-5168 */
  for (int K = 1; K <= KM; K++)
    for (int J = 3; J <= JM - 3; J++)
      for (int I = 1; I <= IM; I++) {
        float ROAVG = 0.25f * (A3(RO, I, J - 2, K) + A3(RO, I, J - 1, K) + A3(RO, I, J + 1, K) + A3(RO, I, J + 2, K));
        if (A3(RO, I, J, K) < 0.6f * ROAVG) A3(RO, I, J, K) = 0.5f * (A3(RO, I, J, K) + 0.6f * ROAVG);
        if (A3(RO, I, J, K) > 1.5f * ROAVG) A3(RO, I, J, K) = 0.5f * (A3(RO, I, J, K) + 1.5f * ROAVG);
        if (A3(RO, I, J, K) < o->ROLIM) A3(RO, I, J, K) = 0.5f * (A3(RO, I, J, K) + o->ROLIM);
      }
  {
    float GM1 = o->GA - 1.0f;
    float TRATIO = 1.0f + o->MACHLIM * o->MACHLIM * 0.5f * GM1;
    float RFGM1 = 1.0f / GM1;
    float RORATIO = powf(TRATIO, RFGM1);
    for (int NR = 1; NR <= o->NROWS; NR++) {
      int JREF = A1(JLE, NR);
      float TREF = A3(T_STATIC, IMID, JREF, KMID);
      float WREFSQ = A3(VX, IMID, JREF, KMID) * A3(VX, IMID, JREF, KMID) +
                     A3(VR, IMID, JREF, KMID) * A3(VR, IMID, JREF, KMID) +
                     A3(WT, IMID, JREF, KMID) * A3(WT, IMID, JREF, KMID);
      float TOREL_REF = TREF + 0.5f * WREFSQ / o->CP;
      float TLIMIT = TOREL_REF / TRATIO;
      if (TLIMIT > o->TCOOL_MIN) TLIMIT = o->TCOOL_MIN;
      A1(VLIMIT, NR) = sqrtf(2 * o->CP * (TOREL_REF - TLIMIT));
      float ROSTAG_REL = A3(RO, IMID, JREF, KMID) * powf(TOREL_REF / TREF, RFGM1);
      A1(ROLIMIT, NR) = ROSTAG_REL / RORATIO;
    }
    for (int K = 1; K <= KM; K++)
      for (int J = 1; J <= JM; J++) {
        int NR = A1(NROW, J);
        float V_LIMIT = A1(VLIMIT, NR), RO_LIMIT = A1(ROLIMIT, NR);
        for (int I = 1; I <= IM; I++) {
          if (A3(RO, I, J, K) < RO_LIMIT) A3(RO, I, J, K) = RO_LIMIT;
          float RONOW = A3(RO, I, J, K);
          float WRELSQ = A3(VX, I, J, K) * A3(VX, I, J, K) + A3(VR, I, J, K) * A3(VR, I, J, K) +
                         A3(WT, I, J, K) * A3(WT, I, J, K);
          float WREL = sqrtf(WRELSQ);
          float FACSAFE = V_LIMIT / WREL;
          if (FACSAFE < 1.0f) {
            A3(ROVX, I, J, K) = RONOW * A3(VX, I, J, K) * FACSAFE;
            A3(ROVR, I, J, K) = RONOW * A3(VR, I, J, K) * FACSAFE;
            float ROVTREL = RONOW * A3(WT, I, J, K) * FACSAFE;
            A3(ROWT, I, J, K) = ROVTREL;
            A3(ROVT, I, J, K) = ROVTREL + RONOW * AJK(UBLADE, J, K);
            A3(RORVT, I, J, K) = A3(ROVT, I, J, K) * AJK(R, J, K);
          }
        }
      }
  }
  if (o->IFMIX > 0 && o->RFMIX > 1.0e-3f) orc_new_mixplan(o);

  /* the trailing-edge separation force of rows with IF_CUSP = 2, mf:5184-5233 */
  for (int NR = 1; o->IF_CUSP && NR <= o->NROWS; NR++) {
    if (A1(IF_CUSP, NR) != 2) continue;
    const int JTEDGE = A1(JTE, NR), JSEP_I1 = JTEDGE - A1(NUP_I1, NR), JSEP_IM = JTEDGE - A1(NUP_IM, NR);
    const int JTHIK = JSEP_I1 > JSEP_IM ? JSEP_I1 : JSEP_IM, NWKE = A1(N_WAKE, NR);
    const float FDRAG = A1(SEP_DRAG, NR);
    for (int K = 1; K <= KM; K++) {
      const float THIK_LIM = A1(SEP_THIK, NR) * AJK(RT_THICK, JTHIK, K);
      const float SSSLOPE = (A3(RTHETA, 1, JSEP_I1 + 1, K) - A3(RTHETA, 1, JSEP_I1 - 1, K)) /
                            (AJK(X, JSEP_I1 + 1, K) - AJK(X, JSEP_I1 - 1, K));
      const float PSSLOPE = (A3(RTHETA, IM, JSEP_IM + 1, K) - A3(RTHETA, IM, JSEP_IM - 1, K)) /
                            (AJK(X, JSEP_IM + 1, K) - AJK(X, JSEP_IM - 1, K));
      for (int side = 0; side < 2; side++) {
        const int JS = side == 0 ? JSEP_I1 : JSEP_IM, IB = side == 0 ? 1 : IM;
        const int I1 = side == 0 ? 1 : IM - 6, I2 = side == 0 ? 6 : IM;
        const float SLOPE = side == 0 ? SSSLOPE : PSSLOPE;
        for (int J = JS; J <= JTEDGE + NWKE; J++) {
          const float RTEXTRAP = A3(RTHETA, IB, JS, K) + SLOPE * (AJK(X, J, K) - AJK(X, JS, K));
          for (int I = I1; I <= I2; I++) {
            const float d = side == 0 ? RTEXTRAP - A3(RTHETA, I, J, K) : A3(RTHETA, I, J, K) - RTEXTRAP;
            if (d > THIK_LIM) {
              A3(ROVX, I, J, K) = FDRAG * A3(ROVX, I, J, K);
              A3(ROVR, I, J, K) = FDRAG * A3(ROVR, I, J, K);
              const float RO_VBLADE = A3(RO, I, J, K) * AJK(UBLADE, J, K);
              const float RO_WT = A3(ROVT, I, J, K) - RO_VBLADE;
              const float RO_VT = FDRAG * RO_WT + RO_VBLADE;
              A3(RORVT, I, J, K) = RO_VT * AJK(R, J, K);
            }
          }
        }
      }
    }
  }

  /* every 5 steps: convergence checks and the time-step update, mf:5252-5349 */
  if (o->NSTEP % 5 == 0 || o->IFEND == 1) {
    for (int J = 1; J <= JM; J++) {
      float SUMAS = 0;
      int NB = A1(NBLADE, J);
      for (int I = 1; I <= IMM1; I++)
        for (int K = 1; K <= KMM1; K++) SUMAS = SUMAS - A3(FLOWX, I, J, K);
      o->BLADE_FLOW[J] = SUMAS * (float)NB;
      o->CHECK_FLOW[J] = (SUMAS + 0.0f) * (float)NB + 0.0f - 0.0f;
    }
    float ECONT = 0.0f;
    for (int J = 1; J <= JM; J++) {
      float RATIO = o->CHECK_FLOW[J] / o->CHECK_FLOW[1];
      float EMASS = fabsf(1.f - RATIO);
      if (EMASS > ECONT) ECONT = EMASS;
    }
    o->ECONT = ECONT;
    o->FLOWRAT = o->BLADE_FLOW[1];
    float EMAX = 0.0f, EAVG = 0.0f;
    float RIJKM = 2.f / (float)(IM * JM * KM);
    float VREF = sqrtf(o->VRMS * RIJKM);
    float RVEF = 100.f / VREF;
    for (int K = 1; K <= KM; K++)
      for (int J = 2; J <= JM; J++) {
        float XD = AJK(DX, J, K), RD = AJK(DR, J, K), SD = AJK(DS, J, K);
        for (int I = 1; I <= IM; I++) {
          float VM_START = (XD * A3(VX, I, J, K) + RD * A3(VR, I, J, K)) / SD;
          float VM_END = (XD * A3(ROVX, I, J, K) + RD * A3(ROVR, I, J, K)) / SD / A3(RO, I, J, K);
          float DVMER = (VM_END - VM_START) * RVEF;
          A3(STORE, I, J, K) = DVMER;
          EAVG = EAVG + fabsf(DVMER);
          if (fabsf(DVMER) > fabsf(EMAX)) EMAX = DVMER;
        }
      }
    o->EAVG = EAVG / (float)(IM * JMM1 * KM);
    o->EMAX = EMAX;
    o->VREF = VREF;
    /* the repeating-stage condition, mf:5323-5331 (N = NSTEP) */
    if (o->IF_REPEAT > 0 && o->NSTEP % o->NINMOD == 0) orc_repeating_stage(o);
    if (o->IFPRINT == 1 || o->IFEND == 1) {
      if (o->ILOS == 10) orc_loss(o);
      if (o->ILOS >= 100 && o->ILOS < 200) orc_new_loss(o);
      if (o->ILOS >= 200) orc_spal_loss(o);
    }
    orc_stepup(o, 0.25f);
  }
}

/* MASS_AVG (mf:11729-11760): the positive-mass-flow average of the node
   field F over each cell line K of plane J, moved to the nodes by the
   running average and the end extrapolations; the plane average. */
static void mass_avg(orc_t *o, int J, const float *F, float *SUMFUN, float *AVGALL) {
  const int IM = o->IM, KM = o->KM;
#define F3(i, j, k) F[((i) - 1) + (size_t)IM * (((j) - 1) + (size_t)o->JM * ((k) - 1))]
  float SUMALL = 0.0f, FLOWALL = 0.0f;
  for (int K = 1; K <= KM - 1; K++) {
    float SUMAS = 0.0f;
    SUMFUN[K] = 0.0f;
    for (int I = 1; I <= IM - 1; I++) {
      float DFLOW = -A3(FLOWX, I, J, K);
      if (DFLOW < 0.0f) DFLOW = 0.0f;
      SUMAS = SUMAS + DFLOW;
      SUMFUN[K] = SUMFUN[K] + DFLOW * (F3(I, J, K) + F3(I + 1, J, K) + F3(I + 1, J, K + 1) + F3(I, J, K + 1)) * 0.25f;
    }
    SUMALL = SUMALL + SUMFUN[K];
    FLOWALL = FLOWALL + SUMAS;
    SUMFUN[K] = SUMFUN[K] / SUMAS;
  }
#undef F3
  for (int K = 2; K <= KM - 1; K++) SUMFUN[K] = 0.5f * (SUMFUN[K] + SUMFUN[K - 1]);
  SUMFUN[1] = 2.0f * SUMFUN[2] - SUMFUN[3];
  SUMFUN[KM] = 2.0f * SUMFUN[KM - 1] - SUMFUN[KM - 2];
  *AVGALL = SUMALL / FLOWALL;
}

/* MIX_BCONDS(1) (mf:11906-12080) over the inlet and every row's JMIX, then
   NEWBCONDS (mf:12085-12115) from the last plane's profiles.  STORE holds
   each node quantity on the plane, as in the reference. */
static void orc_repeating_stage(orc_t *o) {
  const int IM = o->IM, KM = o->KM;
  const float PI = 3.14159265f, DEGRAD = PI / 180.f, RADDEG = 180.f / PI;
  float *po = calloc(KM + 1, sizeof(float)), *f = calloc(KM + 1, sizeof(float));
  float *vt = calloc(KM + 1, sizeof(float)), *yaw = calloc(KM + 1, sizeof(float));
  float avg;
  for (int NR = 1; NR <= o->NROWS + 1; NR++) {
    const int J = NR == 1 ? 1 : A1(JMIX, NR - 1);
    for (int I = 1; I <= IM; I++)
      for (int K = 1; K <= KM; K++) {
        const float VX = A3(VX, I, J, K), VT = A3(VT, I, J, K), VR = A3(VR, I, J, K);
        const float D = .5f * (VX * VX + VT * VT + VR * VR);
        const float TSTAG = A3(HO, I, J, K) / o->CP;
        float TS = TSTAG - D / o->CP;
        if (TS < 1.f) TS = 1.f;
        A3(STORE, I, J, K) = A3(P, I, J, K) * powf(TSTAG / TS, o->GA / (o->GA - 1));
      }
    mass_avg(o, J, o->STORE, po, &avg);
    mass_avg(o, J, o->P, f, &avg);
    for (int I = 1; I <= IM; I++)
      for (int K = 1; K <= KM; K++) A3(STORE, I, J, K) = A3(HO, I, J, K) / o->CP;
    mass_avg(o, J, o->STORE, f, &avg);
    for (int I = 1; I <= IM; I++)
      for (int K = 1; K <= KM; K++) {
        const float VX = A3(VX, I, J, K), VT = A3(VT, I, J, K), VR = A3(VR, I, J, K);
        const float D = .5f * (VX * VX + VT * VT + VR * VR);
        A3(STORE, I, J, K) = (A3(HO, I, J, K) - D) / o->CP;
      }
    mass_avg(o, J, o->STORE, f, &avg);
    mass_avg(o, J, o->VT, vt, &avg);
    for (int I = 1; I <= IM; I++)
      for (int K = 1; K <= KM; K++)
        A3(STORE, I, J, K) = sqrtf(A3(VX, I, J, K) * A3(VX, I, J, K) + A3(VR, I, J, K) * A3(VR, I, J, K));
    mass_avg(o, J, o->STORE, f, &avg);
    for (int I = 1; I <= IM; I++)
      for (int K = 1; K <= KM; K++) {
        const float VMER = sqrtf(A3(VX, I, J, K) * A3(VX, I, J, K) + A3(VR, I, J, K) * A3(VR, I, J, K));
        A3(STORE, I, J, K) = atan2f(A3(VT, I, J, K), VMER) * RADDEG;
      }
    mass_avg(o, J, o->STORE, yaw, &avg);
    for (int I = 1; I <= IM; I++)
      for (int K = 1; K <= KM; K++) {
        const float VMER = sqrtf(A3(VX, I, J, K) * A3(VX, I, J, K) + A3(VR, I, J, K) * A3(VR, I, J, K));
        A3(STORE, I, J, K) = atan2f(A3(VR, I, J, K), VMER) * RADDEG;
      }
    mass_avg(o, J, o->STORE, f, &avg);
  }
  const float RFINBC1 = 1.0f - o->RFINBC;
  const int KMID = o->KMID;
  for (int K = 1; K <= KM; K++) {
    A1(PO1, K) = RFINBC1 * A1(PO1, K) + o->RFINBC * (o->PO_IN_MID + po[K] - po[KMID]);
    A1(BS, K) = RFINBC1 * A1(BS, K) + o->RFINBC * yaw[K];
    A1(VTIN, K) = RFINBC1 * A1(VTIN, K) + o->RFINBC * (o->VT_IN_MID + vt[K] - vt[KMID]);
    A1(RPO1, K) = 1.f / A1(PO1, K);
    A1(BSSIN, K) = sinf(A1(BS, K) * DEGRAD);
    A1(BSCOS, K) = cosf(A1(BS, K) * DEGRAD);
    A1(BRSIN, K) = sinf(A1(BR, K) * DEGRAD);
    A1(BRCOS, K) = cosf(A1(BR, K) * DEGRAD);
  }
  free(po); free(f); free(vt); free(yaw);
}

int orc_step(orc_t *o, int n) {
  int done = 0;
  for (int s = 0; s < n; s++) {
    if (o->IFEND) break;
    step_once(o);
    done++;
  }
  return done;
}
