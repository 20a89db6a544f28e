// C ABI and step orchestration of libmultall_hip (see include/multall_hip.h).
//
// mh_step mirrors the control flow of the reference's LOOP
// (multall-open-18.3.f:3753-5454): host-side bookkeeping of the step counter,
// the smoothing/damping ramp and the stop tests, and one stream of kernel
// launches per step in the reference's phase order.  The every-5-step
// convergence figures come back to the host (one small copy) because the
// next step's stop test needs them, exactly when the reference evaluates it.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <hip/hip_ext.h>

#include "mh_internal.h"

namespace mh {

// kernels (mh_kernels.hip)
__global__ void k_prim(Geo, Dev, Params);
__global__ void k_massflux(Geo, Dev);
__global__ void k_flux_energy(Geo, Dev);
__global__ void k_flux_axial(Geo, Dev);
__global__ void k_flux_angular(Geo, Dev);
__global__ void k_flux_radial(Geo, Dev);
__global__ void k_bc_lines(Geo, Dev, Params);
__global__ void k_bc_pd(Geo, Dev, Params, StepScalars);
__global__ void k_bc_apply(Geo, Dev, Params, StepScalars);
__global__ void k_peff(Geo, Dev, Params, StepScalars);
__global__ void k_fextrap(Geo, Dev, Params);
__global__ void k_residual(Geo, Dev, float *, float *, Params);
__global__ void k_mpavg(Geo, Dev, Params, int);
__global__ void k_mg_sum_seq(Geo, Dev, const int *, const int *, int, int, int, int, int, float *);
__global__ void k_mg_sum_wave(Geo, Dev, const int *, const int *, int, int, int, int, int, float *);
__global__ void k_jplane_sums(Geo, Dev, double *, int);
__global__ void k_range_sums(Geo, const double *, const int *, const int *, int, float *, int, const int *);
__global__ void k_range_sums_seq(Geo, Dev, const int *, const int *, int, float *, int, int);
__global__ void k_mg_apply(Geo, Dev);
__global__ void k_smooth_resid(Geo, Dev, const float *, float *, float, int, int);
__global__ void k_damp_apply(Geo, Dev);
__global__ void k_scatter(Geo, Dev, const float *, float *);
__global__ void k_smooth_plane(Geo, Dev, EqSet, Params, StepScalars, int);
#define SMOOTH_REG_DECL(L, NW, SFX) \
  __global__ void k_smooth_reg##L##w##NW##SFX(Geo, Dev, EqSet, Params, StepScalars, int);
#define SMOOTH_REG_DECLS(SFX)  \
  SMOOTH_REG_DECL(1, 8, SFX)   \
  SMOOTH_REG_DECL(2, 8, SFX)   \
  SMOOTH_REG_DECL(4, 8, SFX)   \
  SMOOTH_REG_DECL(8, 8, SFX)   \
  SMOOTH_REG_DECL(16, 8, SFX)  \
  SMOOTH_REG_DECL(1, 16, SFX)  \
  SMOOTH_REG_DECL(2, 16, SFX)  \
  SMOOTH_REG_DECL(4, 16, SFX)  \
  SMOOTH_REG_DECL(8, 16, SFX)
SMOOTH_REG_DECLS()
SMOOTH_REG_DECLS(c2)
#undef SMOOTH_REG_DECLS
#undef SMOOTH_REG_DECL
__global__ void k_fill_int(int *, int, int);
__global__ void k_rolim_check(Geo, Dev, StepScalars);
__global__ void k_rolim_fix(Geo, Dev, StepScalars, Params);
__global__ void k_mach_apply(Geo, Dev, Params, StepScalars);
__global__ void k_mixplan(Geo, Dev, Params, float *);
__global__ void k_te_sep(Geo, Dev);
__global__ void k_mix_bconds(Geo, Dev, Params);
__global__ void k_newbconds(Geo, Dev, Params);
__global__ void k_vrms_part(Geo, Dev, double *);
__global__ void k_vrms_final(Geo, Dev, const double *, int, int);
__global__ void k_checkflow(Geo, Dev, int);
__global__ void k_econt(Geo, Dev);
__global__ void k_emax_part(Geo, Dev, double *);
__global__ void k_emax_final(Geo, Dev, const double *, int, int);
__global__ void k_stepup_temp(Geo, Dev, Params);
__global__ void k_stepup_step(Geo, Dev, float);
__global__ void k_stepup_mix(Geo, Dev);
__global__ void k_integrals(Geo, Dev, Params);
__global__ void k_eficool(Geo, Dev, Params, const int *, int, double *);
__global__ void k_loss_vislam(Geo, Dev, Params);
__global__ void k_loss_span(Geo, Dev, Params, float);
__global__ void k_loss_pitch(Geo, Dev, Params, float);
__global__ void k_sa_cellavg(Geo, Dev);
__global__ void k_sa_fluxes(Geo, Dev);
__global__ void k_resid_line(Geo, Dev, EqSet, Params, MgFast, int, int, int);
__global__ void k_resid_line_bf(Geo, Dev, EqSet, Params, MgFast, int, int, int);
__global__ void k_resid_line_bfs(Geo, Dev, EqSet, Params, MgFast, int, int, int);
__global__ void k_mg_apply_lines_bfs(Geo, Dev, EqSet);
__global__ void k_sa_tail1_bfs(Geo, Dev, Params, const float *, int, unsigned short *);
__global__ void k_to_bf16(const float *, unsigned short *, unsigned);
__global__ void k_from_bf16(const unsigned short *, float *, unsigned);
__global__ void k_mg_levels(Geo, MgFast, EqSet, const int *, const int *, const int *, const int *);
__global__ void k_mg_apply_lines(Geo, Dev, EqSet);
__global__ void k_flux_mom(Geo, Dev, EqSet, Params, StepScalars);
__global__ void k_massflux_eq(Geo, Dev, EqSet, int);

__global__ void k_xl_dist(Geo, Dev, const float *, float *, int);
__global__ void k_xl_limit(Geo, Dev, const float *, const float *, float *);
__global__ void k_xl_cells(Geo, Dev, const float *, const float *, float, float, float);
__global__ void k_row_avg(Geo, Dev, EqSet, const int *, const int *);
#define DAMP_SCATTER_DECL(NC, TK)                                                  \
  __global__ void k_damp_scatter##NC##_k##TK(Geo, Dev, EqSet, Params, int, int);   \
  __global__ void k_damp_scatter##NC##_k##TK##_bfs(Geo, Dev, EqSet, Params, int, int);
DAMP_SCATTER_DECL(1, 2)
DAMP_SCATTER_DECL(1, 4)
DAMP_SCATTER_DECL(1, 8)
DAMP_SCATTER_DECL(2, 2)
DAMP_SCATTER_DECL(2, 4)
DAMP_SCATTER_DECL(2, 8)
DAMP_SCATTER_DECL(4, 2)
DAMP_SCATTER_DECL(4, 4)
DAMP_SCATTER_DECL(4, 8)
#undef DAMP_SCATTER_DECL
__global__ void k_sa_kvis(Geo, Dev, Params);
__global__ void k_sa_cellavg_kvis(Geo, Dev, Params);
__global__ void k_sa_tail1(Geo, Dev, Params, const float *, int);
__global__ void k_sa_tail2(Geo, Dev, Params, float *);
__global__ void k_sa_grad(Geo, Dev, Params);
__global__ void k_sa_visc(Geo, Dev, Params, int);
__global__ void k_nl_visc(Geo, Dev, Params);
__global__ void k_nl_trans(Geo, Dev, Params);
__global__ void k_pwallgrad_cell(Geo, Dev, Params);
__global__ void k_pwallgrad_smooth(Geo, Dev);
__global__ void k_sa_factors(Geo, Dev, Params);
__global__ void k_setflo_flow(Geo, Dev, int);
__global__ void k_setflo_delta(Geo, Dev, Params, int);
__global__ void k_setflo_apply(Geo, Dev);
__global__ void k_sa_gradtvis(Geo, Dev, Params);
__global__ void k_sa_trcol(Geo, Dev, Params);
__global__ void k_sa_trline(Geo, Dev, Params);
__global__ void k_sa_source(Geo, Dev, Params, float);
__global__ void k_sa_iforces(Geo, Dev, Params);
__global__ void k_sa_kwalls(Geo, Dev, Params);
__global__ void k_sa_kfaces(Geo, Dev, Params, float);
__global__ void k_step_begin(Geo, Dev, Params);
__global__ void k_nan_check(Geo, Dev, Params);

const ArrayInfo kArrays[MH_NARRAYS] = {
    {"RO", K_N, false}, {"ROVX", K_N, false}, {"ROVR", K_N, false}, {"RORVT", K_N, false}, {"ROE", K_N, false},
    {"DRO", K_N, false}, {"DROVX", K_N, false}, {"DROVR", K_N, false}, {"DRORVT", K_N, false}, {"DROE", K_N, false},
    {"ROVT", K_N, false}, {"ROWT", K_N, false}, {"P", K_N, false}, {"HO", K_N, false}, {"PEFF", K_N, false},
    {"VX", K_N, false}, {"VR", K_N, false}, {"VT", K_N, false}, {"WT", K_N, false}, {"T_STATIC", K_N, false},
    {"AQX", K_N, false}, {"AQR", K_N, false}, {"ABX", K_N, false}, {"ABR", K_N, false}, {"ASX", K_N, false},
    {"ASR", K_N, false}, {"VOL", K_N, false}, {"VOLOR", K_N, false}, {"STEP", K_N, false}, {"BSTEP", K_N, false},
    {"RSTEP", K_N, false}, {"XFORCE", K_N, false}, {"RFORCE", K_N, false}, {"TFORCE", K_N, false},
    {"QSOURCE", K_N, false}, {"FLOWX", K_N, false}, {"FLOWT", K_N, false}, {"FLOWR", K_N, false},
    {"R", K_JK, false}, {"X", K_JK, false}, {"RAVG_CELL", K_JK, false}, {"ABT", K_JK, false},
    {"WRABT", K_JK, false}, {"UBLADE", K_JK, false}, {"DX", K_JK, false}, {"DR", K_JK, false},
    {"DS", K_JK, false}, {"SMERID", K_JK, false},
    {"FTL", K_IK, false}, {"FBL", K_IK, false}, {"FTR", K_IK, false}, {"FBR", K_IK, false},
    {"FP", K_I, false}, {"FU", K_I, false}, {"FD", K_I, false},
    {"FR", K_K, false}, {"FKU", K_K, false}, {"FKD", K_K, false}, {"PO1", K_K, false}, {"TO1", K_K, false},
    {"RPO1", K_K, false}, {"BSSIN", K_K, false}, {"BSCOS", K_K, false}, {"BRSIN", K_K, false},
    {"BRCOS", K_K, false}, {"VTIN", K_K, false}, {"PD", K_K, false}, {"VM1", K_K, false},
    {"FACUP", K_J, false}, {"FACDWN", K_J, false}, {"WRAD", K_J, false},
    {"STEP1", K_B1, false}, {"STEP2", K_B2, false}, {"STEPSBK", K_SB, false},
    {"IB1", K_I, true}, {"IB2", K_I, true}, {"KB1", K_K, true}, {"KB2", K_K, true},
    {"NBLADE", K_J, true}, {"IND", K_J, true}, {"INDLE", K_J, true}, {"INDMIX", K_J, true}, {"INDTE", K_J, true},
    {"NROW", K_J, true}, {"JSBLK", K_J, true}, {"JB1", K_J, true}, {"JB2", K_J, true},
    {"JSTART", K_R, true}, {"JMIX", K_R, true}, {"JLE", K_R, true}, {"JTE", K_R, true}, {"JLED", K_R, true},
    {"KTIPS", K_R, true}, {"KTIPE", K_R, true}, {"KS1", K_R, true}, {"KS2", K_R, true},
    {"DWALLSQ", K_N, false}, {"Y_PLUS", K_N, false}, {"WHUB", K_J, false}, {"WTIP", K_J, false},
    {"XLLIM_I1", K_RN, false}, {"XLLIM_IM", K_RN, false}, {"XLLIM_K1", K_RN, false}, {"XLLIM_KM", K_RN, false},
    {"XLLIM_DWN", K_RN, false}, {"XLLIM_UP", K_RN, false}, {"ROUGH_H", K_RN, false}, {"ROUGH_T", K_RN, false},
    {"ROUGH_L", K_RN, false}, {"ROUGH_U", K_RN, false}, {"JTRAN_I1", K_RN, true}, {"JTRAN_IM", K_RN, true},
    {"JTRAN_K1", K_RN, true}, {"JTRAN_KM", K_RN, true},
    {"TRANS_DYN_VIS", K_N, false}, {"DEL_DYNVIS", K_N, false}, {"TRANS_KVIS", K_N, false},
    {"T_SOURCE", K_N, false}, {"ROAVG_CELL", K_N, false}, {"VISC_LAM", K_N, false}, {"VISC_RAT", K_N, false},
    {"XLENGTH", K_N, false}, {"FST_RAT", K_N, false}, {"FSTURB", K_RN, false}, {"TURBVIS_DAMP", K_RN, false},
    {"DPDS_CELL", K_N, false}, {"YPLUS_I1", K_JK, false}, {"YPLUS_IM", K_JK, false}, {"YPLUS_K1", K_IJ, false},
    {"YPLUS_KM", K_IJ, false}, {"CHORD", K_RN, false}, {"NOUT", K_NOUT, true},
    {"NSTAGE", K_RN, true},
    {"IF_CUSP", K_RN, true}, {"NUP_I1", K_RN, true}, {"NUP_IM", K_RN, true}, {"N_WAKE", K_RN, true},
    {"SEP_THIK", K_RN, false}, {"SEP_DRAG", K_RN, false}, {"RT_THICK", K_JK, false}, {"RTHETA", K_N, false},
    {"BS", K_K, false}, {"BR", K_K, false},
};

const char *const kScalarNames[MH_NSCALARS] = {
    "CFL", "DAMPIN", "MACHLIM", "F_PDOWN", "CONLIM", "SFXIN", "SFTIN", "FAC_4TH", "F1", "F2", "F3", "RSMTH",
    "SFEXIT", "FSMTHB", "RFIN", "PDOWN_HUB", "PDOWN_TIP", "FEXTRAP", "RFMIX", "FANGLE", "CP", "GA", "RGAS", "CV",
    "FGA", "GA1", "RCP", "VSOUND", "FACMIXUP", "IMID", "KMID", "NSTEPS_MAX", "NCHANGE", "NRSMTH", "NSFEXIT",
    "IN_PRESS", "IN_VTAN", "IN_VR", "IPOUT", "IFMIX", "IR", "JR", "KR", "ILOS", "NLOS", "NMIXUP", "ITIMST",
    "IF_RESTART", "IFGAS", "IN_FLOW", "IF_REPEAT", "IRBB", "KRBB", "REYNO", "RF_VIS", "FTRANS", "PRANDTL",
    "YPLUSWALL", "YPLAM", "YPTURB", "CFWALL", "IBOUND", "FAC_STMIX", "FAC_ST0", "FAC_ST1", "FAC_ST2", "FAC_ST3",
    "FAC_SFVIS", "FAC_VORT", "FAC_PGRAD", "VORT_MAX", "TURBVIS_LIM", "FLOWIN", "RFLOW", "NINMOD", "RFINBC",
    "PO_IN_MID", "VT_IN_MID", "VISLAM", "FTCOND",
    "EXACT_ORDER", "USE_GRAPHS", "LANES", "MIXED", "FLOW_CONV", "TUNE", "SCAT_TK", "SMOOTH_WAVES",
};

}  // namespace mh

using namespace mh;

// The scratch array of mh_set_xlength_noalloc (see scratch_address below).
constexpr size_t kXlRowSlots = 4 * 64;
constexpr size_t kXlScratchFloats = 2 * (size_t)MH_XLENGTH_SCRATCH_NODES + kXlRowSlots + MH_XLENGTH_SCRATCH_J;
static __device__ __align__(256) float scratch[kXlScratchFloats];

namespace {

constexpr int kRedBlocks = 1024;

// Scratch of mh_set_xlength_noalloc, one array per device, held for the
// life of the process: the nodal THETA and wall distance (NODES floats
// each), the four per-row limits (kXlRowSlots = 4 x 64 rows) and the per-J
// limit (J floats), about 32 MiB.  Host code never passes the array's name
// to a kernel (it is a host-side stand-in there): it takes the device
// address hipGetSymbolAddress gives, per device, at mh_create.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch_ptr[kMaxDevices] = {};

// This device's address of `scratch`, looked up once per device (with that
// device current, as it is on the calling thread).
int scratch_address(int dev, float **out) {
  if (dev < 0 || dev >= kMaxDevices) return MH_E_ARG;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MH_E_HIP;
    g_scratch_ptr[dev] = (float *)p;
  }
  *out = g_scratch_ptr[dev];
  return MH_OK;
}
constexpr size_t kMaxLds = 160 * 1024;  // LDS per workgroup on gfx950

// The dynamic-LDS limit is a per-function, process-wide attribute.  Each
// kernel's limit only grows: a context raises it to what it launches with,
// and a later context with smaller planes never lowers it below what an
// earlier one still needs.
std::mutex g_lds_mu;
std::vector<std::pair<const void *, size_t>> g_lds_limit;

int raise_lds_limit(const void *fn, size_t bytes) {
  std::lock_guard<std::mutex> lk(g_lds_mu);
  for (auto &e : g_lds_limit)
    if (e.first == fn) {
      if (bytes <= e.second) return 0;
      if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess) return 1;
      e.second = bytes;
      return 0;
    }
  if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess) return 1;
  g_lds_limit.push_back({fn, bytes});
  return 0;
}

long long kind_len(const mh_dims &d, Kind k) {
  switch (k) {
    case K_N: return (long long)d.im * d.jm * d.km;
    case K_JK: return (long long)d.jm * d.km;
    case K_IK: return (long long)d.im * d.km;
    case K_IJ: return (long long)d.im * d.jm;
    case K_I: return d.im;
    case K_K: return d.km;
    case K_J: return d.jm;
    case K_R: return d.nrows + 1;
    case K_B1: return (long long)d.nib1 * (d.njb1 + 1) * d.nkb1;
    case K_B2: return (long long)d.nib2 * (d.njb2 + 1) * d.nkb2;
    case K_SB: return d.nsblk;
    case K_RN: return d.nrows;
    case K_NOUT: return 5;
  }
  return 0;
}

void **slot_of(Dev &d, int id) {
  void **base[MH_NARRAYS] = {
      (void **)&d.RO, (void **)&d.ROVX, (void **)&d.ROVR, (void **)&d.RORVT, (void **)&d.ROE, (void **)&d.DRO,
      (void **)&d.DROVX, (void **)&d.DROVR, (void **)&d.DRORVT, (void **)&d.DROE, (void **)&d.ROVT,
      (void **)&d.ROWT, (void **)&d.P, (void **)&d.HO, (void **)&d.PEFF, (void **)&d.VX, (void **)&d.VR,
      (void **)&d.VT, (void **)&d.WT, (void **)&d.T_STATIC, (void **)&d.AQX, (void **)&d.AQR, (void **)&d.ABX,
      (void **)&d.ABR, (void **)&d.ASX, (void **)&d.ASR, (void **)&d.VOL, (void **)&d.VOLOR, (void **)&d.STEP,
      (void **)&d.BSTEP, (void **)&d.RSTEP, (void **)&d.XFORCE, (void **)&d.RFORCE, (void **)&d.TFORCE,
      (void **)&d.QSOURCE, (void **)&d.FLOWX, (void **)&d.FLOWT, (void **)&d.FLOWR, (void **)&d.R, (void **)&d.X,
      (void **)&d.RAVG_CELL, (void **)&d.ABT, (void **)&d.WRABT, (void **)&d.UBLADE, (void **)&d.DX,
      (void **)&d.DR, (void **)&d.DS, (void **)&d.SMERID, (void **)&d.FTL, (void **)&d.FBL, (void **)&d.FTR,
      (void **)&d.FBR, (void **)&d.FP, (void **)&d.FU, (void **)&d.FD, (void **)&d.FR, (void **)&d.FKU,
      (void **)&d.FKD, (void **)&d.PO1, (void **)&d.TO1, (void **)&d.RPO1, (void **)&d.BSSIN, (void **)&d.BSCOS,
      (void **)&d.BRSIN, (void **)&d.BRCOS, (void **)&d.VTIN, (void **)&d.PD, (void **)&d.VM1, (void **)&d.FACUP,
      (void **)&d.FACDWN, (void **)&d.WRAD, (void **)&d.STEP1, (void **)&d.STEP2, (void **)&d.STEPSBK,
      (void **)&d.IB1, (void **)&d.IB2, (void **)&d.KB1, (void **)&d.KB2, (void **)&d.NBLADE, (void **)&d.IND,
      (void **)&d.INDLE, (void **)&d.INDMIX, (void **)&d.INDTE, (void **)&d.NROW, (void **)&d.JSBLK,
      (void **)&d.JB1, (void **)&d.JB2, (void **)&d.JSTART, (void **)&d.JMIX, (void **)&d.JLE, (void **)&d.JTE,
      (void **)&d.JLED, (void **)&d.KTIPS, (void **)&d.KTIPE, (void **)&d.KS1, (void **)&d.KS2,
      (void **)&d.DWALLSQ, (void **)&d.Y_PLUS, (void **)&d.WHUB, (void **)&d.WTIP, (void **)&d.XLLIM_I1,
      (void **)&d.XLLIM_IM, (void **)&d.XLLIM_K1, (void **)&d.XLLIM_KM, (void **)&d.XLLIM_DWN,
      (void **)&d.XLLIM_UP, (void **)&d.ROUGH_H, (void **)&d.ROUGH_T, (void **)&d.ROUGH_L, (void **)&d.ROUGH_U,
      (void **)&d.JTRAN_I1, (void **)&d.JTRAN_IM, (void **)&d.JTRAN_K1, (void **)&d.JTRAN_KM,
      (void **)&d.TRANS_DYN_VIS, (void **)&d.DEL_DYNVIS, (void **)&d.TRANS_KVIS, (void **)&d.T_SOURCE,
      (void **)&d.ROAVG_CELL, (void **)&d.VISC_LAM, (void **)&d.VISC_RAT, (void **)&d.XLENGTH, (void **)&d.FST_RAT,
      (void **)&d.FSTURB, (void **)&d.TURBVIS_DAMP, (void **)&d.DPDS_CELL, (void **)&d.YPLUS_I1,
      (void **)&d.YPLUS_IM, (void **)&d.YPLUS_K1, (void **)&d.YPLUS_KM, (void **)&d.CHORD, (void **)&d.NOUT,
      (void **)&d.NSTAGE, (void **)&d.IF_CUSP, (void **)&d.NUP_I1, (void **)&d.NUP_IM, (void **)&d.N_WAKE,
      (void **)&d.SEP_THIK, (void **)&d.SEP_DRAG, (void **)&d.RT_THICK, (void **)&d.RTHETA, (void **)&d.BS,
      (void **)&d.BR,
  };
  return base[id];
}

int fail(mh_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(call)                                                                 \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) return fail(c, MH_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
int dalloc(mh_ctx *c, T **p, size_t n) {
  void *q = nullptr;
  HIPCHK(hipMalloc(&q, n * sizeof(T) + 256));
  c->owned.push_back(q);
  HIPCHK(hipMemsetAsync(q, 0, n * sizeof(T) + 256, c->stream));
  *p = (T *)q;
  return MH_OK;
}

// All host<->device traffic goes through the context's stream, so it is
// ordered after the zero-fills and kernels already queued there.
int h2d(mh_ctx *c, void *dst, const void *src, size_t bytes) {
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MH_OK;
}
int d2h(mh_ctx *c, void *dst, const void *src, size_t bytes) {
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MH_OK;
}

float sf(const mh_ctx *c, int id) { return (float)c->scal[id]; }
int si(const mh_ctx *c, int id) { return (int)c->scal[id]; }

Params params_of(const mh_ctx *c) {
  Params p{};
  p.CFL = sf(c, MH_S_CFL); p.DAMPIN = sf(c, MH_S_DAMPIN); p.MACHLIM = sf(c, MH_S_MACHLIM);
  p.F_PDOWN = sf(c, MH_S_F_PDOWN); p.CONLIM = sf(c, MH_S_CONLIM); p.SFXIN = c->SFXIN_s; p.SFTIN = c->SFTIN_s;
  p.FAC_4TH = sf(c, MH_S_FAC_4TH); p.F1 = sf(c, MH_S_F1); p.F2 = sf(c, MH_S_F2); p.F3 = sf(c, MH_S_F3);
  p.RSMTH = sf(c, MH_S_RSMTH); p.SFEXIT = sf(c, MH_S_SFEXIT); p.FSMTHB = sf(c, MH_S_FSMTHB);
  p.RFIN = sf(c, MH_S_RFIN); p.PDOWN_HUB = sf(c, MH_S_PDOWN_HUB); p.PDOWN_TIP = sf(c, MH_S_PDOWN_TIP);
  p.FEXTRAP = sf(c, MH_S_FEXTRAP); p.RFMIX = sf(c, MH_S_RFMIX); p.FANGLE = sf(c, MH_S_FANGLE);
  p.CP = sf(c, MH_S_CP); p.GA = sf(c, MH_S_GA); p.RGAS = sf(c, MH_S_RGAS); p.CV = sf(c, MH_S_CV);
  p.FGA = sf(c, MH_S_FGA); p.GA1 = sf(c, MH_S_GA1); p.RCP = sf(c, MH_S_RCP); p.VSOUND = sf(c, MH_S_VSOUND);
  p.FACMIXUP = sf(c, MH_S_FACMIXUP);
  p.IMID = si(c, MH_S_IMID); p.KMID = si(c, MH_S_KMID); p.NSTEPS_MAX = si(c, MH_S_NSTEPS_MAX);
  p.NCHANGE = si(c, MH_S_NCHANGE); p.NRSMTH = si(c, MH_S_NRSMTH); p.NSFEXIT = si(c, MH_S_NSFEXIT);
  p.IN_PRESS = si(c, MH_S_IN_PRESS); p.IN_VTAN = si(c, MH_S_IN_VTAN); p.IN_VR = si(c, MH_S_IN_VR);
  p.IPOUT = si(c, MH_S_IPOUT); p.IFMIX = si(c, MH_S_IFMIX); p.IR = si(c, MH_S_IR); p.JR = si(c, MH_S_JR);
  p.KR = si(c, MH_S_KR); p.ILOS = si(c, MH_S_ILOS); p.NLOS = si(c, MH_S_NLOS); p.NMIXUP = si(c, MH_S_NMIXUP);
  p.ITIMST = si(c, MH_S_ITIMST); p.IF_RESTART = si(c, MH_S_IF_RESTART); p.IFGAS = si(c, MH_S_IFGAS);
  p.IN_FLOW = si(c, MH_S_IN_FLOW); p.IF_REPEAT = si(c, MH_S_IF_REPEAT); p.IRBB = si(c, MH_S_IRBB);
  p.KRBB = si(c, MH_S_KRBB); p.EXACT = si(c, MH_S_EXACT_ORDER);
  p.REYNO = sf(c, MH_S_REYNO); p.RF_VIS = sf(c, MH_S_RF_VIS); p.FTRANS = sf(c, MH_S_FTRANS);
  p.PRANDTL = sf(c, MH_S_PRANDTL); p.YPLUSWALL = sf(c, MH_S_YPLUSWALL); p.YPLAM = sf(c, MH_S_YPLAM);
  p.YPTURB = sf(c, MH_S_YPTURB); p.CFWALL = sf(c, MH_S_CFWALL); p.IBOUND = si(c, MH_S_IBOUND);
  p.FAC_STMIX = sf(c, MH_S_FAC_STMIX); p.FAC_ST0 = sf(c, MH_S_FAC_ST0); p.FAC_ST1 = sf(c, MH_S_FAC_ST1);
  p.FAC_ST2 = sf(c, MH_S_FAC_ST2); p.FAC_ST3 = sf(c, MH_S_FAC_ST3); p.FAC_SFVIS = sf(c, MH_S_FAC_SFVIS);
  p.FAC_VORT = sf(c, MH_S_FAC_VORT); p.FAC_PGRAD = sf(c, MH_S_FAC_PGRAD); p.VORT_MAX = sf(c, MH_S_VORT_MAX);
  p.TURBVIS_LIM = sf(c, MH_S_TURBVIS_LIM);
  p.FLOWIN = sf(c, MH_S_FLOWIN); p.RFLOW = sf(c, MH_S_RFLOW);
  p.NINMOD = si(c, MH_S_NINMOD); p.RFINBC = sf(c, MH_S_RFINBC);
  p.PO_IN_MID = sf(c, MH_S_PO_IN_MID); p.VT_IN_MID = sf(c, MH_S_VT_IN_MID);
  p.FLOW_CONV = sf(c, MH_S_FLOW_CONV);
  return p;
}

}  // namespace

namespace {

dim3 node_grid(const Geo &g) { return dim3((g.IM + 63) / 64, (g.JM + 3) / 4, g.KM); }
const dim3 kNodeBlock(64, 4, 1);

hipEvent_t prof_event(mh_ctx *c) {
  if (!c->prof_pool.empty()) {
    hipEvent_t e = c->prof_pool.back();
    c->prof_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// When profiling is on (mh_profile_enable), a launch goes through
// hipExtLaunchKernel with a start and a stop event that the dispatch itself
// timestamps, so the recorded time is the kernel's own execution (the same
// begin/end that rocprofv3's kernel trace reports).
mh_prof_rec prof_rec(mh_ctx *c, const char *name) {
  mh_prof_rec rec{};
  rec.name = name;
  rec.t0 = prof_event(c);
  rec.t1 = prof_event(c);
  return rec;
}
#define LAUNCH(kern, grid, block, shmem, stream, ...) LAUNCH_NAMED(kern, #kern, grid, block, shmem, stream, __VA_ARGS__)
#define LAUNCH_NAMED(kern, name, grid, block, shmem, stream, ...)                                  \
  do {                                                                                             \
    if (c->prof_on) {                                                                              \
      mh_prof_rec r_ = prof_rec(c, name);                                                          \
      if (r_.t0 && r_.t1) {                                                                        \
        hipExtLaunchKernelGGL(kern, grid, block, shmem, stream, r_.t0, r_.t1, 0, __VA_ARGS__);   \
        c->prof_pending.push_back(r_);                                                             \
        break;                                                                                     \
      }                                                                                            \
    }                                                                                              \
    hipLaunchKernelGGL(kern, grid, block, shmem, stream, __VA_ARGS__);                             \
  } while (0)

// Reduce finished launch timings into the per-kernel totals.
int prof_collect(mh_ctx *c) {
  if (c->prof_pending.empty()) return MH_OK;
  HIPCHK(hipStreamSynchronize(c->stream));
  for (auto &r : c->prof_pending) {
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, r.t0, r.t1));
    bool found = false;
    for (auto &t : c->prof_tot)
      if (t.first == r.name) {
        t.second.first += 1;
        t.second.second += ms;
        found = true;
        break;
      }
    if (!found) c->prof_tot.push_back({r.name, {1, (double)ms}});
    c->prof_pool.push_back(r.t0);
    c->prof_pool.push_back(r.t1);
  }
  c->prof_pending.clear();
  return MH_OK;
}

int last_launch(mh_ctx *c, const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(c, MH_E_HIP, std::string("launch ") + what + ": " + hipGetErrorString(e));
  return MH_OK;
}
#define LCHK(what)                          \
  do {                                      \
    int r_ = last_launch(c, what);          \
    if (r_ != MH_OK) return r_;             \
  } while (0)

StepScalars step_scalars(const mh_ctx *c) {
  StepScalars s{};
  s.PREFF = c->PREFF;
  s.ROLIM = c->ROLIM; s.TCOOL_MIN = 1.0e6f;
  s.SFEX1 = 1.f - sf(c, MH_S_SFEXIT); s.SFEXH = 0.5f * sf(c, MH_S_SFEXIT);
  s.RFIN1 = 1.0f - sf(c, MH_S_RFIN); s.RFIN2 = 2.f * sf(c, MH_S_RFIN);
  return s;
}

// What the host knows of step n before the device takes it: which launches
// the step makes.  Everything that changes from step to step as a value (the
// ramps, the step count, the stop tests) is on the device; the structure of
// the step depends on n alone, except the convergence stop, which the
// device can only find on a step just after a check (n = 5k+1): there the
// check block is launched gated on the device's IFEND.
struct StepPlan {
  bool check5 = false;   // the every-5-step check runs (mf:5240)
  bool cond = false;     // the check runs only if the device ends the run here
  bool visc = false;     // the viscous forces are refreshed (every NLOS steps)
  bool vislam = false;   // ... re-evaluating the laminar viscosity (NSTEP = 1 or < NSTEPS_MAX/4)
  bool refresh = false;  // the closing / printout viscous refresh (mf:5338-5341)
  bool damp = false;     // DAMP within [2, 100]: TSTEP damps
  bool first = false;    // NSTEP = 1
  bool repeat = false;   // the repeating-stage update runs in the check (NSTEP divisible by NINMOD)
  int code() const {
    return check5 | cond << 1 | visc << 2 | vislam << 3 | refresh << 4 | damp << 5 | first << 6 | repeat << 7;
  }
};

// Plan step n (host mirror of the device's ramp for the damping flag; the
// host's DAMP follows the same FP32 statements as k_step_begin).
StepPlan plan_step(mh_ctx *c, int n) {
  const Params p = params_of(c);
  StepPlan q;
  const bool end_known = n == p.NSTEPS_MAX || (c->stop_req && n % 10 == 0);
  q.check5 = n % 5 == 0 || end_known;
  q.cond = !q.check5 && n > 1 && n % 5 == 1;
  q.visc = p.ILOS != 0 && n % p.NLOS == 0;
  q.vislam = n == 1 || n < p.NSTEPS_MAX / 4;
  bool ifprint = false;
  for (int i = 0; i < 5; i++) ifprint = ifprint || n == c->nout[i];
  q.refresh = (q.check5 && ifprint) || end_known;
  const float FCHANGE = 1.0f - (float)(n - 1) / (float)p.NCHANGE;
  if (!(p.IF_RESTART != 0 && n > 100) && FCHANGE > 0.0f) c->DAMP = p.DAMPIN * (1.f - 0.75f * FCHANGE);
  q.damp = !(c->DAMP < 2.0f || c->DAMP > 100);
  q.first = n == 1;
  q.repeat = p.IF_REPEAT > 0 && p.NINMOD > 0 && (q.check5 || q.cond) && n % p.NINMOD == 0;
  return q;
}

// Build the CSR J-lists on the host from the integer maps.
int build_lists(mh_ctx *c) {
  const mh_dims &dm = c->dims;
  std::vector<int> jb1(dm.jm), jb2(dm.jm), jsb(dm.jm), jstart(dm.nrows + 1), jmix(dm.nrows + 1);
  int r0;
  if ((r0 = d2h(c, jb1.data(), c->d.JB1, dm.jm * 4)) || (r0 = d2h(c, jb2.data(), c->d.JB2, dm.jm * 4)) ||
      (r0 = d2h(c, jsb.data(), c->d.JSBLK, dm.jm * 4)) ||
      (r0 = d2h(c, jstart.data(), c->d.JSTART, (dm.nrows + 1) * 4)) ||
      (r0 = d2h(c, jmix.data(), c->d.JMIX, (dm.nrows + 1) * 4)))
    return r0;
  auto csr = [&](const std::vector<int> &map, int nb, std::vector<int> &ptr, std::vector<int> &list) -> bool {
    ptr.assign(nb + 1, 0);
    list.clear();
    for (int b = 1; b <= nb; b++) {
      ptr[b - 1] = (int)list.size();
      for (int J = 2; J <= dm.jm; J++)
        if (map[J - 1] == b) list.push_back(J);
    }
    ptr[nb] = (int)list.size();
    for (int J = 2; J <= dm.jm; J++)
      if (map[J - 1] < 1 || map[J - 1] > nb) return false;
    return true;
  };
  std::vector<int> p1, l1, p2, l2, ps, ls, pr, lr;
  if (!csr(jb1, dm.njb1 + 1, p1, l1)) return fail(c, MH_E_ARG, "JB1 out of range");
  if (!csr(jb2, dm.njb2 + 1, p2, l2)) return fail(c, MH_E_ARG, "JB2 out of range");
  if (!csr(jsb, dm.nsblk, ps, ls)) return fail(c, MH_E_ARG, "JSBLK out of range");
  pr.assign(dm.nrows + 1, 0);
  for (int nr = 1; nr <= dm.nrows; nr++) {
    pr[nr - 1] = (int)lr.size();
    for (int J = jstart[nr - 1] + 1; J <= jmix[nr - 1] - 1; J++) lr.push_back(J);
    if (jmix[nr - 1] - jstart[nr - 1] < 9) return fail(c, MH_E_UNSUPPORTED, "rows need at least 10 J points");
  }
  pr[dm.nrows] = (int)lr.size();
  auto up = [&](std::vector<int> &v, int **dst) -> int {
    int r = dalloc(c, dst, v.size() + 1);
    if (r) return r;
    return h2d(c, *dst, v.data(), v.size() * 4);
  };
  Lists &L = c->lists;
  int r;
  if ((r = up(p1, &L.b1_ptr)) || (r = up(l1, &L.b1_list)) || (r = up(p2, &L.b2_ptr)) || (r = up(l2, &L.b2_list)) ||
      (r = up(ps, &L.sb_ptr)) || (r = up(ls, &L.sb_list)) || (r = up(pr, &L.row_ptr)) || (r = up(lr, &L.row_list)))
    return r;
  // Fast-mode multigrid sums need every level-1 I (K) block of IR (KR) <= 4
  // cells, inside one level-2 I (K) block.
  std::vector<int> ib1(dm.im), ib2(dm.im), kb1(dm.km), kb2(dm.km);
  if ((r = d2h(c, ib1.data(), c->d.IB1, dm.im * 4)) || (r = d2h(c, ib2.data(), c->d.IB2, dm.im * 4)) ||
      (r = d2h(c, kb1.data(), c->d.KB1, dm.km * 4)) || (r = d2h(c, kb2.data(), c->d.KB2, dm.km * 4)))
    return r;
  auto nest = [](const std::vector<int> &a, const std::vector<int> &b, int ncell, int ir, int na, int nb,
                 std::vector<int> &ptr, std::vector<int> &list) -> bool {
    std::vector<int> m(na + 1, 0);
    for (int x = 1; x <= ncell; x++) {
      const int u = a[x - 1], v = b[x - 1];
      if (u < 1 || u > na || v < 1 || v > nb || u != (x - 1) / ir + 1) return false;
      if (m[u] && m[u] != v) return false;
      m[u] = v;
    }
    ptr.assign(nb + 1, 0);
    list.clear();
    for (int v = 1; v <= nb; v++) {
      ptr[v - 1] = (int)list.size();
      for (int u = 1; u <= na; u++)
        if (m[u] == v) list.push_back(u);
    }
    ptr[nb] = (int)list.size();
    return true;
  };
  const Params pp = params_of(c);
  std::vector<int> i2p, i2l, k2p, k2l;
  L.mgf_ok = pp.IR >= 1 && pp.IR <= 4 && pp.KR >= 1 && pp.KR <= 4 &&
             nest(ib1, ib2, dm.im - 1, pp.IR, dm.nib1, dm.nib2, i2p, i2l) &&
             nest(kb1, kb2, dm.km - 1, pp.KR, dm.nkb1, dm.nkb2, k2p, k2l);
  if (L.mgf_ok) {
    MgFast &m = L.mgf;
    // each level-2 block's members as offsets into P[(ib, J, kb)], in the
    // order K block, J, I block
    std::vector<int> b2p, b2o;
    const int nj2 = dm.njb2 + 1;
    for (int kb2 = 0; kb2 < dm.nkb2; kb2++)
      for (int jb2 = 0; jb2 < nj2; jb2++)
        for (int ib2 = 0; ib2 < dm.nib2; ib2++) {
          b2p.push_back((int)b2o.size());
          for (int kq = k2p[kb2]; kq < k2p[kb2 + 1]; kq++)
            for (int jq = p2[jb2]; jq < p2[jb2 + 1]; jq++)
              for (int iq = i2p[ib2]; iq < i2p[ib2 + 1]; iq++)
                b2o.push_back((i2l[iq] - 1) + dm.nib1 * ((l2[jq] - 1) + dm.jm * (k2l[kq] - 1)));
        }
    b2p.push_back((int)b2o.size());
    int *t[2];
    if ((r = up(b2p, &t[0])) || (r = up(b2o, &t[1])) ||
        (r = dalloc(c, &m.P, (size_t)dm.nib1 * dm.jm * dm.nkb1)) || (r = dalloc(c, &m.PT, (size_t)dm.jm * dm.nkb1)))
      return r;
    // level-1 K blocks past the last cell are never written and stay zero
    if (hipMemsetAsync(m.P, 0, (size_t)dm.nib1 * dm.jm * dm.nkb1 * sizeof(double), c->stream) != hipSuccess ||
        hipMemsetAsync(m.PT, 0, (size_t)dm.jm * dm.nkb1 * sizeof(double), c->stream) != hipSuccess)
      return fail(c, MH_E_HIP, "hipMemsetAsync");
    m.b2_ptr = t[0], m.b2_off = t[1];
  }
  // the fused line kernels take rows as contiguous J ranges joined at JMIX
  std::vector<int> nrow(dm.jm);
  if ((r = d2h(c, nrow.data(), c->d.NROW, dm.jm * 4))) return r;
  std::vector<int> jmp(dm.jm, 0);
  for (int nr = 1; nr < dm.nrows; nr++) {
    const int J = jmix[nr - 1];
    if (J + 2 > dm.jm || nrow[J - 1] != nr || nrow[J] != nr + 1 || nrow[J + 1] != nr + 1)
      return fail(c, MH_E_UNSUPPORTED, "rows not contiguous at a mixing plane");
    for (int q = 0; q < 3; q++) jmp[J - 1 + q] = (q + 1) | (J << 2);
  }
  return h2d(c, c->d.JMP, jmp.data(), jmp.size() * 4);
}

// The Dev a lane's kernels see: the context's arrays with the lane's own
// flux and TSTEP scratch (lane 0: the context's own scratch).
Dev lane_dev(const mh_ctx *c, int lane, int gate) {
  Dev d = c->d;
  d.gate = gate;
  if (lane == 0) return d;
  const Lane &l = c->lanes[lane - 1];
  d.XFLUX = l.XFLUX; d.TFLUX = l.TFLUX; d.RFLUX = l.RFLUX; d.SOURCE = l.SOURCE;
  d.STORE = l.STORE; d.DSC = l.DSC; d.TMP = l.TMP;
  d.B1CHG = l.B1CHG; d.B2CHG = l.B2CHG; d.SBCHG = l.SBCHG; d.AVG = l.AVG; d.RED = l.RED;
  return d;
}

// The mixed form's BF16 copy of a history array (DRO, DROVX, DROVR, DRORVT,
// DROE, DEL_DYNVIS), or null for any other array.
int hist_slot(const mh_ctx *c, const float *a) {
  const float *h[kHist] = {c->d.DRO, c->d.DROVX, c->d.DROVR, c->d.DRORVT, c->d.DROE, c->d.DEL_DYNVIS};
  for (int i = 0; i < kHist; i++)
    if (a == h[i]) return i;
  return -1;
}
unsigned short *hist16(const mh_ctx *c, const float *a) {
  const int i = hist_slot(c, a);
  return i < 0 ? nullptr : c->h16[i];
}

// A variable of a batched TSTEP or flux launch on lane `lane`.
EqLane eq_lane(const mh_ctx *c, int lane, float *D, float *DIFF, int ncall) {
  const Dev d = lane_dev(c, lane, -1);
  const MgFast &m = lane ? c->lanes[lane - 1].mgf : c->lists.mgf;
  EqLane q{};
  q.D = D;
  q.DIFF = DIFF;
  q.DIFF16 = hist16(c, DIFF);
  q.STORE16 = lane ? c->lanes[lane - 1].STORE16 : c->store16;
  q.XFLUX = d.XFLUX; q.TFLUX = d.TFLUX; q.RFLUX = d.RFLUX; q.SOURCE = d.SOURCE;
  q.STORE = d.STORE; q.DSC = d.DSC; q.TMP = d.TMP;
  q.B1CHG = d.B1CHG; q.B2CHG = d.B2CHG; q.SBCHG = d.SBCHG; q.AVG = d.AVG;
  q.RED = d.RED; q.P = m.P; q.PT = m.PT;
  q.ncall = ncall;
  return q;
}
// FP32 history <-> its BF16 copy (synchronous; outside the step path).
int hist_convert(mh_ctx *c, int i, bool to_bf16) {
  float *h[kHist] = {c->d.DRO, c->d.DROVX, c->d.DROVR, c->d.DRORVT, c->d.DROE, c->d.DEL_DYNVIS};
  const unsigned n = (unsigned)((size_t)c->g.IM * c->g.JM * c->g.KM);
  if (to_bf16) LAUNCH(k_to_bf16, dim3((n + 255) / 256), dim3(256), 0, c->stream, (const float *)h[i], c->h16[i], n);
  else LAUNCH(k_from_bf16, dim3((n + 255) / 256), dim3(256), 0, c->stream, (const unsigned short *)c->h16[i], h[i], n);
  LCHK("bf16 convert");
  HIPCHK(hipStreamSynchronize(c->stream));
  return MH_OK;
}

// ---------------------------------------------------------------------------
// Tile-shape variants of the damping/corner-distribution kernel (TK node K
// lines per workgroup) and of SMOOTH_VAR's register form (NW wavefronts per J
// plane, L = the smallest power of two >= KM/NW K lines each).  Every variant
// does the same arithmetic for every node, so all give the same bits; which
// is fastest depends on the grid (how many workgroups the launch has next to
// the 256 CUs, how much of each tile is halo), so autotune() times them.
struct ScatVariant {
  void (*k)(Geo, Dev, EqSet, Params, int, int);
  const char *name;
};
// bfs: STORE read from the mixed form's BF16 copy
const ScatVariant &scat_variant(int imm1, int tk, bool bfs = false) {
#define SV(NC, TK, SFX) {k_damp_scatter##NC##_k##TK##SFX, "k_damp_scatter" #NC "_k" #TK #SFX}
  static const ScatVariant v[2][3][3] = {{{SV(1, 2, ), SV(1, 4, ), SV(1, 8, )},
                                          {SV(2, 2, ), SV(2, 4, ), SV(2, 8, )},
                                          {SV(4, 2, ), SV(4, 4, ), SV(4, 8, )}},
                                         {{SV(1, 2, _bfs), SV(1, 4, _bfs), SV(1, 8, _bfs)},
                                          {SV(2, 2, _bfs), SV(2, 4, _bfs), SV(2, 8, _bfs)},
                                          {SV(4, 2, _bfs), SV(4, 4, _bfs), SV(4, 8, _bfs)}}};
#undef SV
  const int nc = scatter_chunks(imm1);
  return v[bfs ? 1 : 0][nc == 1 ? 0 : nc == 2 ? 1 : 2][tk == 2 ? 0 : tk == 4 ? 1 : 2];
}
dim3 scat_grid(const Geo &g, int tk) { return dim3((g.JM + kScatTJ - 1) / kScatTJ, (g.KM + tk - 1) / tk); }

struct SmoothVariant {
  void (*k)(Geo, Dev, EqSet, Params, StepScalars, int);
  const char *name;
  int nw;
};
// nc: points per lane (1 for IM <= 64, 2 for IM <= 128)
const SmoothVariant &smooth_variant(int km, int nw, int nc) {
#define SV(L, NW, SFX) {k_smooth_reg##L##w##NW##SFX, "k_smooth_reg" #L "w" #NW #SFX, NW}
  static const SmoothVariant w8[2][5] = {{SV(1, 8, ), SV(2, 8, ), SV(4, 8, ), SV(8, 8, ), SV(16, 8, )},
                                         {SV(1, 8, c2), SV(2, 8, c2), SV(4, 8, c2), SV(8, 8, c2), SV(16, 8, c2)}};
  static const SmoothVariant w16[2][4] = {{SV(1, 16, ), SV(2, 16, ), SV(4, 16, ), SV(8, 16, )},
                                          {SV(1, 16, c2), SV(2, 16, c2), SV(4, 16, c2), SV(8, 16, c2)}};
#undef SV
  const int lines = (km + nw - 1) / nw;
  const int li = lines <= 1 ? 0 : lines <= 2 ? 1 : lines <= 4 ? 2 : lines <= 8 ? 3 : 4;
  const int ci = nc > 1 ? 1 : 0;
  return nw == 16 ? w16[ci][std::min(li, 3)] : w8[ci][li];
}
// the register form of SMOOTH_VAR holds IM <= 128 points of a line and
// KM <= 128 lines of a plane
bool smooth_reg_fits(const Geo &g) { return g.IM <= 128 && g.KM <= 128; }
int smooth_chunks(const Geo &g) { return g.IM <= 64 ? 1 : 2; }

// The tuner: for each kernel, one launch to warm up, then three launches
// between two events on the context's stream, for every shape; the fastest
// is kept, per grid shape for the whole process.  The launches read the
// context's own arrays and write only the scratch arrays (the corner
// distribution's output DSC, and TMP in place of SMOOTH_VAR's output), so
// the state is left as it was.
int autotune(mh_ctx *c) {
  static std::mutex mu;
  static std::map<std::array<int, 6>, std::pair<int, int>> known;
  const Geo &g = c->g;
  const Params p = params_of(c);
  const bool exit_sm = !(p.SFEXIT < 0.0001f) && p.NSFEXIT > 0;  // the exit planes' sweeps change the timing
  const std::array<int, 6> key = {g.IM, g.JM, g.KM, g.NROWS, p.EXACT, exit_sm ? 1 : 0};
  const int ftk = (int)c->scal[MH_S_SCAT_TK], fnw = (int)c->scal[MH_S_SMOOTH_WAVES];
  if (ftk && ftk != 2 && ftk != 4 && ftk != 8) return fail(c, MH_E_ARG, "SCAT_TK must be 0, 2, 4 or 8");
  if (fnw && fnw != 8 && fnw != 16) return fail(c, MH_E_ARG, "SMOOTH_WAVES must be 0, 8 or 16");
  if (fnw == 16 && g.KM > 128) return fail(c, MH_E_ARG, "SMOOTH_WAVES = 16 needs KM <= 128");
  c->scat_tk = ftk ? ftk : 8;
  c->smooth_nw = fnw ? fnw : 8;
  if (c->scal[MH_S_TUNE] == 0.0 || (ftk && fnw)) return MH_OK;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = known.find(key);
    if (it != known.end()) {
      if (!ftk) c->scat_tk = it->second.first;
      if (!fnw) c->smooth_nw = it->second.second;
      return MH_OK;
    }
  }
  const bool prof = c->prof_on;
  c->prof_on = false;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  const Dev &d = c->d;
  const StepScalars s = step_scalars(c);
  // the shape the step launches: three variables at once in the fast mode
  // (the lanes' own scratch), one in the exact mode; the corner
  // distribution reads the conserved variables and writes DSC, SMOOTH_VAR
  // reads DSC and writes TMP in place of the variable
  const int nz = c->nlanes == kLanes ? 3 : 1;
  float *const vars[3] = {d.ROVX, d.RORVT, d.ROVR};
  EqSet es{}, et{};
  for (int i = 0; i < kMaxEq; i++) {
    const int l = i < nz ? i : 0;
    es.q[i] = eq_lane(c, l, vars[l], d.DROVX, 4);
    et.q[i] = es.q[i];
    et.q[i].D = et.q[i].TMP;
  }
  auto time_it = [&](auto launch) -> float {
    launch();
    (void)hipEventRecord(e0, c->stream);
    for (int i = 0; i < 3; i++) launch();
    (void)hipEventRecord(e1, c->stream);
    float ms = 0.0f;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1e30f;
    return ms;
  };
  int best_tk = c->scat_tk, best_nw = c->smooth_nw;
  if (!ftk) {
    float best = 1e30f;
    for (int tk : kScatTKs) {
      const ScatVariant &v = scat_variant(g.IMM1, tk);
      const float t = time_it([&] {
        dim3 sg = scat_grid(g, tk);
        sg.z = nz;
        hipLaunchKernelGGL(v.k, sg, dim3(256), scatter_lds_bytes(g.IMM1, tk), c->stream, g, d, es, p, 1, 0);
      });
      if (t < best) best = t, best_tk = tk;
    }
  }
  if (!fnw && smooth_reg_fits(g)) {
    float best = 1e30f;
    for (int nw : kSmoothWaves) {
      const SmoothVariant &v = smooth_variant(g.KM, nw, smooth_chunks(g));
      const float t = time_it([&] {
        hipLaunchKernelGGL(v.k, dim3(g.JM, 1, nz), dim3(64 * v.nw), 0, c->stream, g, d, et, p, s, 1);
      });
      if (t < best) best = t, best_nw = nw;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->prof_on = prof;
  HIPCHK(hipGetLastError());
  c->scat_tk = best_tk;
  c->smooth_nw = best_nw;
  std::lock_guard<std::mutex> lk(mu);
  known[key] = {best_tk, best_nw};
  return MH_OK;
}

// One TSTEP call (mf:5470-5932) for each of n variables whose time steps do
// not read each other's results (continuity, energy and the turbulence
// transport; the three momentum equations), on the context's stream, gated
// on CTL slot `gate`.  The chain: the fused line kernel (flux extrapolation,
// residual, first mixing-plane average, multigrid partial sums), the
// multigrid sums, the multigrid update (with the damping line sums in fast
// mode), the row averages, the fused damping + second mixing-plane average +
// corner distribution, and SMOOTH_VAR with D's boundary conditions.  Each
// variable has its own flux and TSTEP scratch (lane lanes[i]); the hot
// kernels take all n in one launch (grid z), the others run per variable.
// A transport variable (ncall = 2) comes last: it leaves after the damping
// averages for its own smoothing and conditions.
int tstep(mh_ctx *c, const EqLane *eqs, const int *lanes, int n, const StepPlan &q, int gate) {
  const Geo &g = c->g;
  const hipStream_t st = c->stream;
  const Params p = params_of(c);
  const StepScalars s = step_scalars(c);
  const dim3 grid = node_grid(g);
  const dim3 lblock(256);
  const int nmp = (g.NROWS - 1) * g.KMM1;
  const Lists &L = c->lists;
  if (n < 1 || n > kMaxEq) return fail(c, MH_E_ARG, "tstep: 1 to 3 variables");
  EqSet e{};
  for (int i = 0; i < kMaxEq; i++) e.q[i] = eqs[i < n ? i : 0];
  const Dev d0 = lane_dev(c, 0, gate);
  auto dev_of = [&](int i) { return lane_dev(c, lanes[i], gate); };
  MgFast mgf = L.mgf;
  mgf.gate = c->d.CTL + gate;
  const int *gp = c->d.CTL + gate;
  const bool damp = q.damp;
  // fast mode without residual smoothing takes the damping line sums from
  // the multigrid update
  const bool fused_apply = !p.EXACT && p.NRSMTH == 0 && damp;
  const bool mg = !(p.IR <= 1 && p.JR <= 1 && p.KR <= 1);
  // fast mode: the residual kernel also leaves the multigrid partial sums
  const bool fastmg = mg && !p.EXACT && L.mgf_ok;
  const int kg = fastmg ? p.KR : 4;
  // the mixed form keeps STORE in BF16 through the fused fast-mode chain
  // (residual, multigrid update with the damping line sums, damping and
  // corner distribution); the other chains read it as FP32
  bool bfs = c->mixed && fused_apply && fastmg;
  for (int i = 0; i < n; i++) bfs = bfs && eqs[i].STORE16 != nullptr;
  const dim3 rgrid((g.JM - 1 + 3) / 4, (g.KMM1 + kg - 1) / kg, n);
  if (bfs) LAUNCH(k_resid_line_bfs, rgrid, lblock, 0, st, g, d0, e, p, mgf, kg, p.IR, 1);
  else if (c->mixed) LAUNCH(k_resid_line_bf, rgrid, lblock, 0, st, g, d0, e, p, mgf, kg, p.IR, fastmg ? 1 : 0);
  else LAUNCH(k_resid_line, rgrid, lblock, 0, st, g, d0, e, p, mgf, kg, p.IR, fastmg ? 1 : 0);
  LCHK("resid_line");
  if (mg) {
    const int n1 = g.NIB1 * (g.NJB1 + 1) * g.NKB1, n2 = g.NIB2 * (g.NJB2 + 1) * g.NKB2;
    if (fastmg) {
      LAUNCH(k_mg_levels, dim3((n1 + 255) / 256 + (n2 + 3) / 4 + g.NSBLK, n), dim3(256), 0, st, g, mgf, e,
             L.b1_ptr, L.b1_list, L.sb_ptr, L.sb_list);
    } else {
      for (int i = 0; i < n; i++) {
        const Dev d = dev_of(i);
        if (p.EXACT) {
          LAUNCH(k_mg_sum_seq, dim3((n1 + 127) / 128), dim3(128), 0, st, g, d, L.b1_ptr, L.b1_list, g.NIB1,
                 g.NJB1 + 1, g.NKB1, p.IR, p.KR, d.B1CHG);
          LAUNCH(k_mg_sum_seq, dim3((n2 + 63) / 64), dim3(64), 0, st, g, d, L.b2_ptr, L.b2_list, g.NIB2,
                 g.NJB2 + 1, g.NKB2, p.IRBB, p.KRBB, d.B2CHG);
          LAUNCH(k_range_sums_seq, dim3((g.NSBLK + 63) / 64), dim3(64), 0, st, g, d, L.sb_ptr, L.sb_list,
                 g.NSBLK, d.SBCHG, 0, 0);
        } else {
          LAUNCH(k_mg_sum_seq, dim3((n1 + 127) / 128), dim3(128), 0, st, g, d, L.b1_ptr, L.b1_list, g.NIB1,
                 g.NJB1 + 1, g.NKB1, p.IR, p.KR, d.B1CHG);
          LAUNCH(k_mg_sum_wave, dim3(n2), dim3(64), 0, st, g, d, L.b2_ptr, L.b2_list, g.NIB2, g.NJB2 + 1,
                 g.NKB2, p.IRBB, p.KRBB, d.B2CHG);
          LAUNCH(k_jplane_sums, dim3(g.JM), dim3(256), 0, st, g, d, d.RED, 0);
          LAUNCH(k_range_sums, dim3(g.NSBLK), dim3(64), 0, st, g, (const double *)d.RED, L.sb_ptr, L.sb_list,
                 g.NSBLK, d.SBCHG, 0, gp);
        }
      }
    }
    LCHK("mg_sums");
  }
  if (fused_apply) {
    const dim3 ag((g.JM - 1 + 3) / 4, (g.KMM1 + 3) / 4, n);
    if (bfs) LAUNCH(k_mg_apply_lines_bfs, ag, lblock, 0, st, g, d0, e);
    else LAUNCH(k_mg_apply_lines, ag, lblock, 0, st, g, d0, e);
  } else {
    for (int i = 0; i < n; i++) LAUNCH(k_mg_apply, grid, kNodeBlock, 0, st, g, dev_of(i));
  }
  LCHK("mgapply");
  if (p.NRSMTH > 0) {
    // STORE -> TMP -> DSC -> STORE (DSC is free until the corner
    // distribution), so the buffers keep their roles from call to call
    for (int i = 0; i < n; i++) {
      const Dev d = dev_of(i);
      for (int nsm = 0; nsm < p.NRSMTH; nsm++) {
        LAUNCH(k_smooth_resid, grid, kNodeBlock, 0, st, g, d, (const float *)d.STORE, d.TMP, p.RSMTH, 0, 0);
        LAUNCH(k_smooth_resid, grid, kNodeBlock, 0, st, g, d, (const float *)d.TMP, d.DSC, p.RSMTH, 0, 1);
        LAUNCH(k_smooth_resid, grid, kNodeBlock, 0, st, g, d, (const float *)d.DSC, d.STORE, p.RSMTH, 0, 2);
      }
    }
    LCHK("smooth_resid");
  }
  // row averages of |change| for the damping (mf:5677-5711)
  if (fused_apply) {
    LAUNCH(k_row_avg, dim3(g.NROWS, n), dim3(1024), 0, st, g, d0, e, L.row_ptr, L.row_list);
    LCHK("row_avg");
  } else if (damp) {
    for (int i = 0; i < n; i++) {
      const Dev d = dev_of(i);
      if (p.EXACT) {
        LAUNCH(k_range_sums_seq, dim3(1), dim3(64), 0, st, g, d, L.row_ptr, L.row_list, g.NROWS, d.AVG, 1, 1);
      } else {
        LAUNCH(k_jplane_sums, dim3(g.JM), dim3(256), 0, st, g, d, d.RED, 1);
        LAUNCH(k_range_sums, dim3(g.NROWS), dim3(64), 0, st, g, (const double *)d.RED, L.row_ptr, L.row_list,
               g.NROWS, d.AVG, 1, gp);
      }
    }
    LCHK("damp_avg");
  }
  int nv = n;
  while (nv > 0 && eqs[nv - 1].ncall == 2) nv--;
  for (int i = nv; i < n; i++) {
    // the cell-centred turbulent viscosity (mf:5736-5777, then 8700):
    // add the (damped) change, carry the pitch average over each mixing
    // plane, smooth once with FAC_SFVIS*SFT, then the periodic/tip/
    // mixing-plane conditions with ILAST = IM-1
    // (two line kernels: the new value, the pitch average over each mixing
    // plane and the streamwise and pitchwise smoothing into TMP; then the
    // spanwise smoothing and the line conditions back into D; the smoothing
    // factor FAC_SFVIS * SFT with the step's SFT taken on the device)
    const Dev d = dev_of(i);
    float *D = eqs[i].D;
    const dim3 tg((g.KM + 3) / 4, g.JM);
    if (bfs) LAUNCH(k_sa_tail1_bfs, tg, dim3(256), 0, st, g, d, p, (const float *)D, damp ? 1 : 0, eqs[i].STORE16);
    else LAUNCH(k_sa_tail1, tg, dim3(256), 0, st, g, d, p, (const float *)D, damp ? 1 : 0);
    LAUNCH(k_sa_tail2, tg, dim3(256), 0, st, g, d, p, D);
    LCHK("sa_tstep");
  }
  if (nv == 0) return MH_OK;
  // damping, second mixing-plane averages (mf:5786-5818) and the corner
  // distribution
  {
    const ScatVariant &v = scat_variant(g.IMM1, c->scat_tk, bfs);
    dim3 sg = scat_grid(g, c->scat_tk);
    sg.z = nv;
    LAUNCH_NAMED(v.k, v.name, sg, dim3(256), scatter_lds_bytes(g.IMM1, c->scat_tk), st, g, d0, e, p, damp ? 1 : 0,
                 nmp > 0 ? 1 : 0);
  }
  LCHK("damp_scatter");
  // SMOOTH_VAR: planes in registers (exit-smoothing planes included) where
  // IM <= 128 and KM <= 128, the LDS form elsewhere
  if (smooth_reg_fits(g)) {
    const SmoothVariant &v = smooth_variant(g.KM, c->smooth_nw, smooth_chunks(g));
    LAUNCH_NAMED(v.k, v.name, dim3(g.JM, 1, nv), dim3(64 * v.nw), 0, st, g, d0, e, p, s, 1);
  } else {
    LAUNCH(k_smooth_plane, dim3(g.JM, 1, nv), dim3(512), c->lds_plane, st, g, d0, e, p, s, 1);
  }
  LCHK("smooth");
  return MH_OK;
}

// TSTEP for one variable on lane 0
int tstep1(mh_ctx *c, float *D, float *DIFF, int ncall, const StepPlan &q, int gate) {
  const EqLane e = eq_lane(c, 0, D, DIFF, ncall);
  const int lane = 0;
  return tstep(c, &e, &lane, 1, q, gate);
}

int stepup(mh_ctx *c, const Dev &d, float relax) {
  const Geo &g = c->g;
  const Params p = params_of(c);
  LAUNCH(k_stepup_temp, node_grid(g), kNodeBlock, 0, c->stream, g, d, p);
  LAUNCH(k_stepup_step, node_grid(g), kNodeBlock, 0, c->stream, g, d, relax);
  const int nmix = (g.NROWS - 1) * g.IMM1 * g.KMM1;
  if (nmix > 0) LAUNCH(k_stepup_mix, dim3((nmix + 255) / 256), dim3(256), 0, c->stream, g, d);
  LCHK("stepup");
  return MH_OK;
}

// Mixing-length factors of LOSS (mf:8125-8183), formed on the host in FP32
// in the reference's statement order.  They depend on the grid only, so they
// are formed once (the reference recomputes the same values while NSTEP <=
// NLOS).
int loss_tables(mh_ctx *c) {
  const Geo &g = c->g;
  const Dev &d = c->d;
  const int IM = g.IM, KM = g.KM, JM = g.JM, IMM1 = g.IMM1, KMM1 = g.KMM1, NR = g.NROWS, KMID = g.KMID;
  std::vector<float> FP(IM), FR(KM), X((size_t)JM * KM), R((size_t)JM * KM);
  std::vector<float> xi1(NR), xim(NR), xk1(NR), xkm(NR), xdw(NR), xup(NR);
  std::vector<int> NB(JM), JLE(NR + 1), JTE(NR + 1);
  int r;
  if ((r = d2h(c, FP.data(), d.FP, IM * 4)) || (r = d2h(c, FR.data(), d.FR, KM * 4)) ||
      (r = d2h(c, X.data(), d.X, X.size() * 4)) || (r = d2h(c, R.data(), d.R, R.size() * 4)) ||
      (r = d2h(c, NB.data(), d.NBLADE, JM * 4)) || (r = d2h(c, JLE.data(), d.JLE, NR * 4)) ||
      (r = d2h(c, JTE.data(), d.JTE, NR * 4)) || (r = d2h(c, xi1.data(), d.XLLIM_I1, NR * 4)) ||
      (r = d2h(c, xim.data(), d.XLLIM_IM, NR * 4)) || (r = d2h(c, xk1.data(), d.XLLIM_K1, NR * 4)) ||
      (r = d2h(c, xkm.data(), d.XLLIM_KM, NR * 4)) || (r = d2h(c, xdw.data(), d.XLLIM_DWN, NR * 4)) ||
      (r = d2h(c, xup.data(), d.XLLIM_UP, NR * 4)))
    return r;
  std::vector<float> t((size_t)NR * (4 * IM + 2 * KM), 0.0f);
  auto P = [&](int tab, int nr, int I) -> float & { return t[lt_pitch(tab, nr, I, NR, IM)]; };
  auto S = [&](int tab, int nr, int K) -> float & { return t[lt_span(tab, nr, K, NR, IM, KM)]; };
  auto AJK = [&](const std::vector<float> &a, int J, int K) { return a[(size_t)(J - 1) + (size_t)JM * (K - 1)]; };
  for (int nrw = 1; nrw <= NR; nrw++) {
    const float dw = xdw[nrw - 1], up = xup[nrw - 1];
    float SUMF = 0.0f;
    for (int I = 2; I <= IMM1; I++) {
      SUMF = SUMF + FP[I - 2];
      const float XLIM = xi1[nrw - 1] * (1.0f - SUMF) + xim[nrw - 1] * SUMF;
      float FPITCH = SUMF * (1.f - SUMF);
      if (FPITCH > XLIM) FPITCH = XLIM;
      const float DF = FP[I - 1] + FP[I - 2];
      P(LT_FILAM, nrw, I) = FP[I - 1] / DF;
      P(LT_FITURB, nrw, I) = 0.16f * FPITCH * FPITCH / (DF * DF);
      P(LT_FIWAKE, nrw, I) = 0.16f * dw * dw / (DF * DF);
      P(LT_FIUP, nrw, I) = 0.16f * up * up / (DF * DF);
    }
    const float DF = FP[0] + FP[IMM1 - 1];
    P(LT_FILAM, nrw, 1) = FP[0] / DF;
    P(LT_FIWAKE, nrw, 1) = 0.16f * dw * dw / (DF * DF);
    P(LT_FIUP, nrw, 1) = 0.16f * up * up / (DF * DF);
    P(LT_FITURB, nrw, 1) = P(LT_FITURB, nrw, 2);
    P(LT_FILAM, nrw, IM) = FP[IMM1 - 1] / DF;
    P(LT_FIWAKE, nrw, IM) = 0.16f * dw * dw / (DF * DF);
    P(LT_FIUP, nrw, IM) = 0.16f * up * up / (DF * DF);
    P(LT_FITURB, nrw, IM) = P(LT_FITURB, nrw, IMM1);
    float AVGSPAN = 0.0f;
    const int JMID = (int)(0.5f * (float)(JLE[nrw - 1] + JTE[nrw - 1]));
    if (JMID < 1 || JMID > JM) return fail(c, MH_E_ARG, "LOSS: row leading/trailing edge outside the grid");
    for (int K = 2; K <= KM; K++) {
      const float XDIF = AJK(X, JMID, K) - AJK(X, JMID, K - 1);
      const float RDIF = AJK(R, JMID, K) - AJK(R, JMID, K - 1);
      AVGSPAN = AVGSPAN + sqrtf(XDIF * XDIF + RDIF * RDIF);
    }
    const float AVGPIT = 2.0f * 3.1415926f * AJK(R, JMID, KMID) / (float)NB[JMID - 1];
    const float XLIMH = xk1[nrw - 1] * AVGPIT / AVGSPAN;
    const float XLIMT = xkm[nrw - 1] * AVGPIT / AVGSPAN;
    SUMF = 0.0f;
    for (int K = 2; K <= KMM1; K++) {
      SUMF = SUMF + FR[K - 2];
      const float XLIM = XLIMH * (1.0f - SUMF) + XLIMT * SUMF;
      float FCSPAN = SUMF * (1.f - SUMF);
      if (FCSPAN > XLIM) FCSPAN = XLIM;
      const float DF2 = FR[K - 1] + FR[K - 2];
      S(LT_FKLAM, nrw, K) = FR[K - 1] / DF2;
      S(LT_FKTURB, nrw, K) = 0.16f * FCSPAN * FCSPAN / (DF2 * DF2);
    }
  }
  return h2d(c, d.LTAB, t.data(), t.size() * 4);
}

// SET_PWALLGRAD (mf:16052-16114): the smoothed streamwise pressure gradient
// the Shih wall functions (YPLUSWALL < -10) and the SA pressure-gradient
// source factor read.  A pure function of PEFF and the velocities, which a
// viscous call does not change, so one evaluation serves each call.
int pwallgrad(mh_ctx *c, const Dev &d) {
  const Geo &g = c->g;
  const Params p = params_of(c);
  LAUNCH(k_pwallgrad_cell, node_grid(g), kNodeBlock, 0, c->stream, g, d, p);
  LAUNCH(k_pwallgrad_smooth, dim3((g.IMM1 * g.KMM1 + 63) / 64), dim3(64), 0, c->stream, g, d);
  LCHK("pwallgrad");
  return MH_OK;
}

// One LOSS call (mf:8058-9066) with the given relaxation factor; FMIXUP is
// the device's (SC_FMIXUP).  vislam: the laminar viscosity is re-evaluated
// (NSTEP = 1 or NSTEP < NSTEPS_MAX/4, mf:8082).
int loss(mh_ctx *c, const Dev &d, float rf_vis, bool vislam) {
  const Geo &g = c->g;
  const Params p = params_of(c);
  if (vislam) LAUNCH(k_loss_vislam, dim3(1), dim3(64), 0, c->stream, g, d, p);
  int r;
  if (p.YPLUSWALL < -10.0f && (r = pwallgrad(c, d))) return r;  // mf:8197
  LAUNCH(k_loss_span, dim3((g.IMM1 + 63) / 64, (g.JM - 1 + 3) / 4), dim3(64, 4), 0, c->stream, g, d, p, rf_vis);
  LAUNCH(k_loss_pitch, dim3(g.KMM1, g.JM - 1), dim3(64), 0, c->stream, g, d, p, rf_vis);
  LCHK("loss");
  return MH_OK;
}

// Per-J values of the variables the periodic-force block of SPAL_LOSS and
// NEW_LOSS reads without setting them (its J-1 and row number keep the values of the last
// blade J its K, J loop visited; before the first one, those left by its
// earlier loops), and the row its hub/casing loop then reads.
int sa_maps(mh_ctx *c) {
  const Geo &g = c->g;
  const int JM = g.JM;
  std::vector<int> IND(JM), NROW(JM), INDMIX(JM), map(4 * JM + 4, 0);
  int r;
  if ((r = d2h(c, IND.data(), c->d.IND, JM * 4)) || (r = d2h(c, NROW.data(), c->d.NROW, JM * 4)) ||
      (r = d2h(c, INDMIX.data(), c->d.INDMIX, JM * 4)))
    return r;
  if (params_of(c).ILOS >= 200)  // k_sa_visc copies the vorticity from J-1 at a mixing plane
    for (int J = 2; J <= JM; J++)
      if (INDMIX[J - 1] == 1 && INDMIX[J - 2] == 1) return fail(c, MH_E_UNSUPPORTED, "adjacent mixing-plane J");
  int last = 0;
  for (int J = 2; J <= JM; J++)
    if (IND[J - 1] == 1) last = J;
  for (int pass = 0; pass < 2; pass++) {  // K = 1, then K > 1
    int jm1 = JM - 1, nrw = NROW[JM - 1];
    if (pass == 1 && last) { jm1 = last - 1; nrw = NROW[last - 1]; }
    for (int J = 2; J <= JM; J++) {
      if (IND[J - 1] == 1) { jm1 = J - 1; nrw = NROW[J - 1]; }
      map[pass * JM + J - 1] = jm1;
      map[(2 + pass) * JM + J - 1] = nrw;
    }
  }
  map[4 * JM] = last ? NROW[last - 1] : NROW[JM - 1];
  return h2d(c, c->d.SAJMAP, map.data(), map.size() * 4);
}

// One SPAL_LOSS call (mf:13586-14827) with the given relaxation factor;
// first: NSTEP = 1.
int spal_loss(mh_ctx *c, const Dev &d, float rf_vis, bool vislam, bool first) {
  const Geo &g = c->g;
  const Params p = params_of(c);
  const dim3 grid = node_grid(g);
  if (vislam) LAUNCH(k_loss_vislam, dim3(1), dim3(64), 0, c->stream, g, d, p);
  int r;
  // mf:13656-13658 and 14175
  if ((p.FAC_PGRAD > 0.001f || p.YPLUSWALL < -10.0f) && (r = pwallgrad(c, d))) return r;
  LAUNCH(k_sa_grad, grid, kNodeBlock, 0, c->stream, g, d, p);
  LAUNCH(k_sa_visc, grid, kNodeBlock, 0, c->stream, g, d, p, first ? 1 : 0);
  if (p.FAC_VORT > 0.001f || p.FAC_PGRAD > 0.001f) LAUNCH(k_sa_factors, dim3(1), dim3(64), 0, c->stream, g, d, p);
  LAUNCH(k_sa_gradtvis, grid, kNodeBlock, 0, c->stream, g, d, p);
  LAUNCH(k_sa_trcol, dim3((g.IMM1 + 63) / 64, g.JM - 1), dim3(64), 0, c->stream, g, d, p);
  LAUNCH(k_sa_trline, dim3((g.KMM1 + 63) / 64, g.JM - 1), dim3(64), 0, c->stream, g, d, p);
  LAUNCH(k_sa_source, grid, kNodeBlock, 0, c->stream, g, d, p, rf_vis);
  LAUNCH(k_sa_iforces, dim3(g.KMM1, g.JM - 1), dim3(64), 0, c->stream, g, d, p);
  LAUNCH(k_sa_kwalls, dim3((g.IMM1 + 63) / 64, (g.JM - 1 + 3) / 4), dim3(64, 4), 0, c->stream, g, d, p);
  LAUNCH(k_sa_kfaces, dim3((g.IMM1 + 63) / 64, (g.JM - 1 + 3) / 4, (g.KMM1 + kKSeg - 1) / kKSeg), dim3(64, 4), 0,
         c->stream, g, d, p, rf_vis);
  LCHK("spal_loss");
  return MH_OK;
}

// One NEW_LOSS call (mf:12522-13462) with the given relaxation factor:
// SPAL_LOSS's gradient, stress and force kernels around the mixing-length
// viscosity and the transition test.
int new_loss(mh_ctx *c, const Dev &d, float rf_vis, bool vislam) {
  const Geo &g = c->g;
  const Params p = params_of(c);
  const dim3 grid = node_grid(g);
  if (vislam) LAUNCH(k_loss_vislam, dim3(1), dim3(64), 0, c->stream, g, d, p);
  LAUNCH(k_sa_grad, grid, kNodeBlock, 0, c->stream, g, d, p);
  LAUNCH(k_nl_visc, grid, kNodeBlock, 0, c->stream, g, d, p);
  LAUNCH(k_nl_trans, dim3(g.JM - 1), dim3(64), 0, c->stream, g, d, p);
  int r;
  if (p.YPLUSWALL < -10.0f && (r = pwallgrad(c, d))) return r;  // mf:12812
  LAUNCH(k_sa_iforces, dim3(g.KMM1, g.JM - 1), dim3(64), 0, c->stream, g, d, p);
  LAUNCH(k_sa_kwalls, dim3((g.IMM1 + 63) / 64, (g.JM - 1 + 3) / 4), dim3(64, 4), 0, c->stream, g, d, p);
  LAUNCH(k_sa_kfaces, dim3((g.IMM1 + 63) / 64, (g.JM - 1 + 3) / 4, (g.KMM1 + kKSeg - 1) / kKSeg), dim3(64, 4), 0,
         c->stream, g, d, p, rf_vis);
  LCHK("new_loss");
  return MH_OK;
}

// The viscous model's call for ILOS (mf:3935-3942, 5339-5341, 7354-7356).
int viscous(mh_ctx *c, const Dev &d, int ilos, float rf_vis, bool vislam, bool first) {
  if (ilos == 10) return loss(c, d, rf_vis, vislam);
  if (ilos >= 100 && ilos < 200) return new_loss(c, d, rf_vis, vislam);
  if (ilos >= 200) return spal_loss(c, d, rf_vis, vislam, first);
  return MH_OK;
}

// Forget the captured step graphs (after a scalar changes, at a new start,
// at destroy).
void drop_graphs(mh_ctx *c) {
  for (auto &kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
  c->graphs.clear();
}

// The flux and TSTEP scratch of the batched time steps' second and third
// variables (fast mode), made once per context.
static int make_lanes(mh_ctx *c) {
  if (c->nlanes == kLanes) return MH_OK;
  const mh_dims &dm = c->dims;
  const size_t N = (size_t)dm.im * dm.jm * dm.km;
  int r;
  for (int l = 0; l < kLanes; l++) {
    Lane &L = c->lanes[l];
    if ((r = dalloc(c, &L.XFLUX, N)) || (r = dalloc(c, &L.TFLUX, N)) || (r = dalloc(c, &L.RFLUX, N)) ||
        (r = dalloc(c, &L.SOURCE, N)) || (r = dalloc(c, &L.STORE, N)) || (r = dalloc(c, &L.DSC, N)) ||
        (r = dalloc(c, &L.TMP, N)) || (r = dalloc(c, &L.B1CHG, (size_t)kind_len(dm, K_B1))) ||
        (r = dalloc(c, &L.B2CHG, (size_t)kind_len(dm, K_B2))) || (r = dalloc(c, &L.SBCHG, (size_t)dm.nsblk)) ||
        (r = dalloc(c, &L.AVG, (size_t)64)) ||
        (r = dalloc(c, &L.RED, std::max<size_t>(4 * kRedBlocks, (size_t)dm.jm * dm.km))))
      return r;
    if (c->lists.mgf_ok &&
        ((r = dalloc(c, &L.mgf.P, (size_t)dm.nib1 * dm.jm * dm.nkb1)) || (r = dalloc(c, &L.mgf.PT, (size_t)dm.jm * dm.nkb1))))
      return r;
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  c->nlanes = kLanes;
  return MH_OK;
}

// One time step: the body of LOOP from label 5000 (mf:3753-5454), queued on
// the context's streams without waiting for anything.  The step's
// bookkeeping runs on the device (k_step_begin); every launch after it is
// gated on the device's CTL_SKIP (the run has ended), and the check block of
// a step that may end the run by convergence (q.cond) on CTL_NOTEND.
int enqueue_step(mh_ctx *c, const StepPlan &q) {
  const Geo &g = c->g;
  const Params p = params_of(c);
  const dim3 grid = node_grid(g);
  int r;
  const Dev d = lane_dev(c, 0, CTL_SKIP);
  const Dev dn = lane_dev(c, 0, CTL_NOTEND);     // launches only the run's last step makes
  const Dev &dc = q.check5 ? d : dn;               // the check block
  const StepScalars s = step_scalars(c);
  {
    Dev d0 = c->d;
    d0.gate = -1;
    LAUNCH(k_step_begin, dim3(1), dim3(64), 0, c->stream, g, d0, p);
  }
  LAUNCH(k_prim, grid, kNodeBlock, 0, c->stream, g, d, p);
  LCHK("prim");
  if (q.check5 || q.cond) {
    LAUNCH(k_vrms_part, dim3(kRedBlocks), dim3(256), 0, c->stream, g, dc, d.RED);
    LAUNCH(k_vrms_final, dim3(1), dim3(256), 0, c->stream, g, dc, (const double *)d.RED, kRedBlocks, p.EXACT);
    LCHK("vrms");
  }
  // viscous forces every NLOS steps, mf:3935-3942
  if (q.visc && (r = viscous(c, d, p.ILOS, p.RF_VIS, q.vislam, q.first))) return r;
  // Batched (fast mode, LANES = 1): continuity, energy and the turbulence
  // transport take their time steps together (their fluxes read only the
  // mass fluxes, HO, QSOURCE, the previous step's kinematic viscosity and
  // T_SOURCE; only the kinematic value that closes the transport needs the
  // new density), then, after the boundary conditions, the three momentum
  // equations (fluxes from the step's primitives, PEFF and the body forces).
  // Each variable has its own flux arrays and TSTEP scratch (lanes 0-2;
  // continuity's fluxes are the mass-flux arrays; the energy and transport
  // fluxes are formed beside them, in the same launch).
  // Otherwise one variable at a time in the reference's order.
  const bool batched = c->nlanes == kLanes && !p.EXACT && c->scal[MH_S_LANES] != 0.0;
  if (!batched) {
    LAUNCH(k_massflux, grid, kNodeBlock, 0, c->stream, g, d);
    LCHK("massflux");
  }
  if (batched) {
    EqLane eq[kMaxEq];
    int ln[kMaxEq] = {0, 1, 2};
    int n = 0;
    // continuity's fluxes are the mass fluxes themselves, with no source
    eq[n] = eq_lane(c, 0, d.RO, d.DRO, 1);
    eq[n].XFLUX = d.FLOWX;
    eq[n].TFLUX = d.FLOWT;
    eq[n].RFLUX = d.FLOWR;
    eq[n++].SOURCE = nullptr;
    eq[n++] = eq_lane(c, 1, d.ROE, d.DROE, 3);
    if (p.ILOS >= 200) eq[n++] = eq_lane(c, 2, d.TRANS_DYN_VIS, d.DEL_DYNVIS, 2);
    EqSet ef{};
    for (int i = 0; i < kMaxEq; i++) ef.q[i] = eq[std::min(i + 1, n - 1)];
    LAUNCH(k_massflux_eq, grid, kNodeBlock, 0, c->stream, g, d, ef, n - 1);
    LCHK("massflux_eq");
    if ((r = tstep(c, eq, ln, n, q, CTL_SKIP))) return r;
    if (p.ILOS >= 100) {
      // cell-average density (mf:4260-4270), then the transport's lower
      // limit and kinematic value
      if (p.ILOS >= 200) LAUNCH(k_sa_cellavg_kvis, grid, kNodeBlock, 0, c->stream, g, d, p);
      else LAUNCH(k_sa_cellavg, grid, kNodeBlock, 0, c->stream, g, d);
      LCHK("sa_transport");
    }
  } else {
    if ((r = tstep1(c, d.RO, d.DRO, 1, q, CTL_SKIP))) return r;
    if (p.ILOS >= 100 && p.ILOS < 200) {
      // cell-average density for NEW_LOSS, mf:4260-4270
      LAUNCH(k_sa_cellavg, grid, kNodeBlock, 0, c->stream, g, d);
      LCHK("cellavg");
    }
    if (p.ILOS >= 200) {
      // cell-average density (mf:4260-4270), the transport (mf:4275-4335),
      // then the lower limit and kinematic value
      LAUNCH(k_sa_cellavg, grid, kNodeBlock, 0, c->stream, g, d);
      LAUNCH(k_sa_fluxes, grid, kNodeBlock, 0, c->stream, g, d);
      if ((r = tstep1(c, d.TRANS_DYN_VIS, d.DEL_DYNVIS, 2, q, CTL_SKIP))) return r;
      LAUNCH(k_sa_kvis, grid, kNodeBlock, 0, c->stream, g, d, p);
      LCHK("sa_transport");
    }
    LAUNCH(k_flux_energy, grid, kNodeBlock, 0, c->stream, g, d);
    LCHK("flux_energy");
    if ((r = tstep1(c, d.ROE, d.DROE, 3, q, CTL_SKIP))) return r;
  }
  LAUNCH(k_bc_lines, dim3(g.KM), dim3(64), 0, c->stream, g, d, p);
  LAUNCH(k_bc_pd, dim3(1), dim3(128), 0, c->stream, g, d, p, s);
  LAUNCH(k_bc_apply, dim3((g.IM * g.KM + 255) / 256), dim3(256), 0, c->stream, g, d, p, s);
  LCHK("bc");
  if (p.IN_FLOW != 0) {
    // mass-flow forcing, mf:4746
    LAUNCH(k_setflo_flow, dim3(g.JM), dim3(64), 0, c->stream, g, d, p.EXACT);
    LAUNCH(k_setflo_delta, dim3(1), dim3(256), 0, c->stream, g, d, p, p.EXACT);
    LAUNCH(k_setflo_apply, grid, kNodeBlock, 0, c->stream, g, d);
    LCHK("setflo");
  }
  // PEFF (mf:4751-4757): the batched momentum flux kernel forms it itself
  if (!batched) LAUNCH(k_peff, grid, kNodeBlock, 0, c->stream, g, d, p, s);
  if (batched) {
    EqSet em{};
    em.q[0] = eq_lane(c, 0, d.ROVX, d.DROVX, 4);
    em.q[1] = eq_lane(c, 1, d.RORVT, d.DRORVT, 5);
    em.q[2] = eq_lane(c, 2, d.ROVR, d.DROVR, 6);
    const int ln[kMaxEq] = {0, 1, 2};
    LAUNCH(k_flux_mom, grid, kNodeBlock, 0, c->stream, g, d, em, p, s);
    LCHK("flux_mom");
    if ((r = tstep(c, em.q, ln, 3, q, CTL_SKIP))) return r;
  } else {
    LAUNCH(k_flux_axial, grid, kNodeBlock, 0, c->stream, g, d);
    LCHK("flux_axial");
    if ((r = tstep1(c, d.ROVX, d.DROVX, 4, q, CTL_SKIP))) return r;
    LAUNCH(k_flux_angular, grid, kNodeBlock, 0, c->stream, g, d);
    LCHK("flux_angular");
    if ((r = tstep1(c, d.RORVT, d.DRORVT, 5, q, CTL_SKIP))) return r;
    LAUNCH(k_flux_radial, grid, kNodeBlock, 0, c->stream, g, d);
    LCHK("flux_radial");
    if ((r = tstep1(c, d.ROVR, d.DROVR, 6, q, CTL_SKIP))) return r;
  }
  // limiters, mf:5115-5168
  const int nline = g.IM * g.KM;
  LAUNCH(k_rolim_check, grid, kNodeBlock, 0, c->stream, g, d, s);
  LAUNCH(k_rolim_fix, dim3((nline + 63) / 64), dim3(64), 0, c->stream, g, d, s, p);
  LAUNCH(k_mach_apply, grid, kNodeBlock, 0, c->stream, g, d, p, s);
  LCHK("limiters");
  if (g.IM > 2 && p.IFMIX > 0 && p.RFMIX > 1.0e-3f && g.NROWS > 1) {
    const int nl = (g.NROWS - 1) * g.KM;
    LAUNCH(k_mixplan, dim3(nl), dim3(64), 9 * g.IM * sizeof(float), c->stream, g, d, p, c->lists.mixscratch);
    LCHK("mixplan");
  }
  if (c->te_sep) {
    LAUNCH(k_te_sep, dim3(g.NROWS * g.KM), dim3(256), 0, c->stream, g, d);
    LCHK("te_sep");
  }
  if (q.check5 || q.cond) {
    LAUNCH(k_checkflow, dim3(g.JM), dim3(64), 0, c->stream, g, dc, p.EXACT);
    LAUNCH(k_econt, dim3(1), dim3(256), 0, c->stream, g, dc);
    LAUNCH(k_emax_part, dim3(kRedBlocks), dim3(256), 0, c->stream, g, dc, d.RED);
    LAUNCH(k_emax_final, dim3(1), dim3(256), 0, c->stream, g, dc, (const double *)d.RED, kRedBlocks, p.EXACT);
    LAUNCH(k_nan_check, dim3(1), dim3(64), 0, c->stream, g, dc, p);
    LCHK("convergence");
    // the repeating-stage condition, mf:5323-5331: MIX_BCONDS, then NEWBCONDS
    if (q.repeat) {
      LAUNCH(k_mix_bconds, dim3((g.NROWS + 1) * 8), dim3(64), (g.KM + 1) * sizeof(double), c->stream, g, dc, p);
      LAUNCH(k_newbconds, dim3((g.KM + 63) / 64), dim3(64), 0, c->stream, g, dc, p);
      LCHK("repeating stage");
    }
    // forces brought up to date before a printout or the final output,
    // mf:5338-5341 (on a step that may end by convergence: if it does)
    if (q.refresh && (r = viscous(c, dc, p.ILOS, p.RF_VIS, q.vislam, q.first))) return r;
    if (q.cond && !q.refresh && (r = viscous(c, dn, p.ILOS, p.RF_VIS, q.vislam, q.first))) return r;
    if ((r = stepup(c, dc, 0.25f))) return r;
  }
  return MH_OK;
}

// Steps n0+1 .. n0+m as one captured HIP graph (built on first use and kept
// per step pattern), or queued launch by launch when graphs are off or the
// kernels are being timed.
int enqueue_chunk(mh_ctx *c, const std::vector<StepPlan> &plan) {
  const bool graphs = c->scal[MH_S_USE_GRAPHS] != 0.0 && !c->prof_on;
  if (!graphs) {
    for (const StepPlan &q : plan) {
      int r = enqueue_step(c, q);
      if (r) return r;
    }
    return MH_OK;
  }
  std::string key;
  for (const StepPlan &q : plan) key += std::to_string(q.code()) + ",";
  auto it = c->graphs.find(key);
  if (it == c->graphs.end()) {
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int r = MH_OK;
    for (const StepPlan &q : plan)
      if ((r = enqueue_step(c, q))) break;
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &gr);
    if (r) {
      if (gr) (void)hipGraphDestroy(gr);
      return r;
    }
    if (e != hipSuccess || !gr) return fail(c, MH_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    hipGraphExec_t ex = nullptr;
    const hipError_t e2 = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e2 != hipSuccess) return fail(c, MH_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e2));
    it = c->graphs.emplace(key, ex).first;
  }
  HIPCHK(hipGraphLaunch(it->second, c->stream));
  return MH_OK;
}

// The run's state as the device left it: step count, end flag and the last
// check's figures (one small copy).
int read_back(mh_ctx *c) {
  int ctl[CTL_COUNT];
  float sc[SC_NHEAD];
  HIPCHK(hipMemcpyAsync(ctl, c->d.CTL, sizeof ctl, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(sc, c->d.SC, sizeof sc, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->nstep = ctl[CTL_NSTEP];
  c->ifend = (ctl[CTL_IFEND] || ctl[CTL_DONE]) ? 1 : 0;
  c->last.nstep = c->nstep;
  c->last.ifend = c->ifend;
  c->last.emax = sc[SC_EMAX];
  c->last.eavg = sc[SC_EAVG];
  c->last.econt = sc[SC_ECONT];
  c->last.vref = sc[SC_VREF];
  c->last.vmax = sc[SC_VMAX];
  c->last.flow_in = sc[SC_FLOWRAT];
  c->last.flow_out = sc[SC_FLOWOUT];
  c->last.imax = ctl[CTL_IMAX];
  c->last.jmax = ctl[CTL_JMAX];
  c->last.kmax = ctl[CTL_KMAX];
  if (ctl[CTL_NANSTEP])
    return fail(c, MH_E_NAN, "non-finite flow field at step " + std::to_string(ctl[CTL_NANSTEP]));
  return MH_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI

extern "C" {

int mh_array_id(const char *name) {
  if (!name) return -1;
  for (int i = 0; i < MH_NARRAYS; i++)
    if (!strcmp(name, kArrays[i].name)) return i;
  return -1;
}

int mh_scalar_id(const char *name) {
  if (!name) return -1;
  for (int i = 0; i < MH_NSCALARS; i++)
    if (!strcmp(name, kScalarNames[i])) return i;
  return -1;
}

int mh_array_is_int(int id) { return (id < 0 || id >= MH_NARRAYS) ? -1 : (kArrays[id].is_int ? 1 : 0); }

long long mh_array_len(const mh_ctx *c, int id) {
  if (!c || id < 0 || id >= MH_NARRAYS) return -1;
  return kind_len(c->dims, kArrays[id].kind);
}

const char *mh_last_error(const mh_ctx *c) { return c ? c->err.c_str() : "null context"; }

int mh_create(const mh_dims *dims, mh_ctx **out) {
  if (!dims || !out) return MH_E_ARG;
  *out = nullptr;
  const mh_dims &dm = *dims;
  // pitchwise lines are summed on one wavefront in up to 4 chunks (IM-1 <= 256);
  // the exit-pressure recursion stages up to 512 K values in LDS
  if (dm.im < 4 || dm.im > 257 || dm.km < 4 || dm.km > 512 || dm.jm < 10 || dm.nrows < 1 || dm.nrows > 64 || dm.nib1 < 1 || dm.njb1 < 1 ||
      dm.nkb1 < 1 || dm.nib2 < 1 || dm.njb2 < 1 || dm.nkb2 < 1 || dm.nsblk < 1)
    return MH_E_ARG;
  mh_ctx *c = new mh_ctx();
  c->dims = dm;
  if (hipGetDevice(&c->device) != hipSuccess || scratch_address(c->device, &c->xl_scratch) != MH_OK) {
    delete c;
    return MH_E_HIP;
  }
  Geo &g = c->g;
  g.IM = dm.im; g.JM = dm.jm; g.KM = dm.km; g.IMM1 = dm.im - 1; g.JMM1 = dm.jm - 1; g.KMM1 = dm.km - 1;
  g.NROWS = dm.nrows; g.NIB1 = dm.nib1; g.NJB1 = dm.njb1; g.NKB1 = dm.nkb1; g.NIB2 = dm.nib2; g.NJB2 = dm.njb2;
  g.NKB2 = dm.nkb2; g.NSBLK = dm.nsblk;
  auto bail = [&](int r) { mh_destroy(c); return r; };
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return bail(MH_E_HIP);
  int r;
  for (int id = 0; id < MH_NARRAYS; id++) {
    const long long n = kind_len(dm, kArrays[id].kind);
    void **slot = slot_of(c->d, id);
    if (kArrays[id].is_int) r = dalloc(c, (int **)slot, (size_t)n);
    else r = dalloc(c, (float **)slot, (size_t)n);
    if (r) return bail(r);
  }
  const size_t N = (size_t)dm.im * dm.jm * dm.km;
  Dev &d = c->d;
  {
    // the wall functions' y+ state starts at SETUP's 25 (mf:6906-6910)
    const std::vector<float> y25((size_t)std::max(dm.jm * dm.km, dm.im * dm.jm), 25.0f);
    if ((r = h2d(c, d.YPLUS_I1, y25.data(), (size_t)dm.jm * dm.km * 4)) ||
        (r = h2d(c, d.YPLUS_IM, y25.data(), (size_t)dm.jm * dm.km * 4)) ||
        (r = h2d(c, d.YPLUS_K1, y25.data(), (size_t)dm.im * dm.jm * 4)) ||
        (r = h2d(c, d.YPLUS_KM, y25.data(), (size_t)dm.im * dm.jm * 4)))
      return bail(r);
  }
  if ((r = dalloc(c, &d.XFLUX, N)) || (r = dalloc(c, &d.TFLUX, N)) || (r = dalloc(c, &d.RFLUX, N)) ||
      (r = dalloc(c, &d.SOURCE, N)) || (r = dalloc(c, &d.STORE, N)) || (r = dalloc(c, &d.DSC, N)) ||
      (r = dalloc(c, &d.TMP, N)) || (r = dalloc(c, &d.TEMP1, N)) ||
      (r = dalloc(c, &d.SXX, N)) || (r = dalloc(c, &d.SXR, N)) || (r = dalloc(c, &d.SXT, N)) ||
      (r = dalloc(c, &d.SRR, N)) || (r = dalloc(c, &d.SRT, N)) || (r = dalloc(c, &d.STT, N)) ||
      (r = dalloc(c, &d.GQX, N)) || (r = dalloc(c, &d.GQR, N)) || (r = dalloc(c, &d.GQT, N)) ||
      (r = dalloc(c, &d.ABS_VORT, N)) || (r = dalloc(c, &d.VISC_TURB, N)) || (r = dalloc(c, &d.DYN_VIS_ML, N)) ||
      (r = dalloc(c, &d.GRAD_TVIS, N)) || (r = dalloc(c, &d.TVGX, N)) || (r = dalloc(c, &d.TVGR, N)) ||
      (r = dalloc(c, &d.TVGT, N)) || (r = dalloc(c, &d.FORCEX, N)) || (r = dalloc(c, &d.FORCER, N)) ||
      (r = dalloc(c, &d.FORCET, N)) || (r = dalloc(c, &d.ESOURCE, N)) ||
      (r = dalloc(c, &d.B1CHG, kind_len(dm, K_B1))) || (r = dalloc(c, &d.B2CHG, kind_len(dm, K_B2))) ||
      (r = dalloc(c, &d.SBCHG, (size_t)dm.nsblk)) || (r = dalloc(c, &d.BCW, (size_t)4 * dm.km + 4)) ||
      (r = dalloc(c, &d.LTAB, (size_t)dm.nrows * (4 * dm.im + 2 * dm.km))) ||
      (r = dalloc(c, &d.SAFLAG, (size_t)dm.im * dm.jm + (size_t)dm.jm * dm.km)) ||
      (r = dalloc(c, &d.SAJMAP, (size_t)4 * dm.jm + 4)) || (r = dalloc(c, &d.JMP, (size_t)dm.jm)) ||
      (r = dalloc(c, &d.RED, std::max<size_t>(4 * kRedBlocks, (size_t)dm.jm * dm.km))) ||
      (r = dalloc(c, &d.SC, (size_t)SC_COUNT + dm.km)) || (r = dalloc(c, &d.FLAGS, (size_t)dm.im * dm.km)) ||
      (r = dalloc(c, &d.CHECK_FLOW, (size_t)dm.jm)) || (r = dalloc(c, &d.BLADE_FLOW, (size_t)dm.jm)) ||
      (r = dalloc(c, &d.SETFLO, (size_t)2 * dm.jm)) ||
      (r = dalloc(c, &d.KWALL, (size_t)kKWall * dm.im * dm.jm)) ||
      (r = dalloc(c, &d.INTEG, (size_t)7 * dm.jm)) || (r = dalloc(c, &d.CTL, (size_t)CTL_COUNT)) ||
      (r = dalloc(c, &d.MIXB, (size_t)(dm.nrows + 1) * 8 * (dm.km + 1))) ||
      (r = dalloc(c, &c->lists.mixscratch, (size_t)std::max(1, dm.nrows - 1) * dm.km * 9 * dm.im)))
    return bail(r);
  d.AVG = d.SC + SC_AVG_CHG;
  d.gate = -1;
  c->scal[MH_S_USE_GRAPHS] = 1.0;  // default: steps run as captured graphs
  c->scal[MH_S_LANES] = 1.0;       // default: concurrent equation streams (fast mode)
  c->scal[MH_S_TUNE] = 1.0;        // default: tile shapes timed at mh_start
  if (hipHostMalloc((void **)&c->ctl_host, CTL_COUNT * sizeof(int), hipHostMallocDefault) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_poll, hipEventDisableTiming) != hipSuccess)
    return bail(MH_E_HIP);
  for (auto &e : c->ev_ahead)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return bail(MH_E_HIP);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return bail(MH_E_HIP);
  c->lds_plane = ((size_t)3 * dm.im * dm.km + 2 * (size_t)(dm.im + dm.km)) * sizeof(float);
  if (c->lds_plane > kMaxLds || scatter_lds_bytes(dm.im - 1, 8) > kMaxLds) return bail(MH_E_UNSUPPORTED);
  if (raise_lds_limit((const void *)k_smooth_plane, c->lds_plane))
    return bail(fail(c, MH_E_HIP, "hipFuncSetAttribute(MaxDynamicSharedMemorySize)"));
  for (int tk : kScatTKs)
    if (raise_lds_limit((const void *)scat_variant(dm.im - 1, tk).k, scatter_lds_bytes(dm.im - 1, tk)))
      return bail(fail(c, MH_E_HIP, "hipFuncSetAttribute(MaxDynamicSharedMemorySize)"));
  *out = c;
  return MH_OK;
}

int mh_destroy(mh_ctx *c) {
  if (!c) return MH_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto &r : c->prof_pending) {
    (void)hipEventDestroy(r.t0);
    (void)hipEventDestroy(r.t1);
  }
  for (auto e : c->prof_pool) (void)hipEventDestroy(e);
  drop_graphs(c);
  if (c->ev_poll) (void)hipEventDestroy(c->ev_poll);
  for (auto e : c->ev_ahead)
    if (e) (void)hipEventDestroy(e);
  if (c->ctl_host) (void)hipHostFree(c->ctl_host);
  for (void *p : c->owned) (void)hipFree(p);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return MH_OK;
}

static int copy_array_raw(mh_ctx *c, int id, void *host, long long n, int ld_i, int ld_j, bool to_dev);

// In the mixed form the library's copy of a history array is its BF16
// array: reads widen it first, writes narrow it after.
static int copy_array(mh_ctx *c, int id, void *host, long long n, int ld_i, int ld_j, bool to_dev) {
  if (!c || id < 0 || id >= MH_NARRAYS || !host) return fail(c, MH_E_ARG, "bad argument");
  const int h = (c->mixed && kArrays[id].kind == K_N) ? hist_slot(c, (const float *)*slot_of(c->d, id)) : -1;
  int r;
  if (h >= 0 && !to_dev && (r = hist_convert(c, h, false))) return r;
  if ((r = copy_array_raw(c, id, host, n, ld_i, ld_j, to_dev))) return r;
  if (h >= 0 && to_dev && (r = hist_convert(c, h, true))) return r;
  return MH_OK;
}

static int copy_array_raw(mh_ctx *c, int id, void *host, long long n, int ld_i, int ld_j, bool to_dev) {
  const mh_dims &dm = c->dims;
  const Kind k = kArrays[id].kind;
  const long long len = kind_len(dm, k);
  char *dev = (char *)*slot_of(c->d, id);
  const hipMemcpyKind dir = to_dev ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
  if (ld_i <= 0 && ld_j <= 0) {  // dense
    if (n != len) return fail(c, MH_E_ARG, std::string("size mismatch for ") + kArrays[id].name);
    return to_dev ? h2d(c, dev, host, (size_t)len * 4) : d2h(c, host, dev, (size_t)len * 4);
  }
  // padded FORTRAN array: rows of the first logical extent at pitch ld_i
  size_t w = 0, h = 0, planes = 1, spitch = 0, splane = 0;
  switch (k) {
    case K_N: w = dm.im; h = dm.jm; planes = dm.km; spitch = ld_i; splane = (size_t)ld_i * ld_j; break;
    case K_B1: w = dm.nib1; h = dm.njb1 + 1; planes = dm.nkb1; spitch = ld_i; splane = (size_t)ld_i * ld_j; break;
    case K_B2: w = dm.nib2; h = dm.njb2 + 1; planes = dm.nkb2; spitch = ld_i; splane = (size_t)ld_i * ld_j; break;
    case K_JK: w = dm.jm; h = dm.km; spitch = ld_j; break;
    case K_IK: w = dm.im; h = dm.km; spitch = ld_i; break;
    case K_IJ: w = dm.im; h = dm.jm; spitch = ld_i; break;
    default: w = (size_t)len; h = 1; spitch = w; break;
  }
  if (spitch < w || (planes > 1 && (size_t)ld_j < h)) return fail(c, MH_E_ARG, "leading dimensions too small");
  for (size_t pz = 0; pz < planes; pz++) {
    char *dptr = dev + pz * w * h * 4;
    char *hptr = (char *)host + pz * splane * 4;
    if (to_dev) HIPCHK(hipMemcpy2DAsync(dptr, w * 4, hptr, spitch * 4, w * 4, h, dir, c->stream));
    else HIPCHK(hipMemcpy2DAsync(hptr, spitch * 4, dptr, w * 4, w * 4, h, dir, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return MH_OK;
}

int mh_set_array(mh_ctx *c, int id, const void *host, long long n) {
  return copy_array(c, id, const_cast<void *>(host), n, 0, 0, true);
}
int mh_get_array(mh_ctx *c, int id, void *host, long long n) { return copy_array(c, id, host, n, 0, 0, false); }
int mh_set_array_strided(mh_ctx *c, int id, const void *host, int ld_i, int ld_j) {
  return copy_array(c, id, const_cast<void *>(host), 0, ld_i > 0 ? ld_i : 1, ld_j > 0 ? ld_j : 1, true);
}
int mh_get_array_strided(mh_ctx *c, int id, void *host, int ld_i, int ld_j) {
  return copy_array(c, id, host, 0, ld_i > 0 ? ld_i : 1, ld_j > 0 ? ld_j : 1, false);
}

int mh_set_scalar(mh_ctx *c, int id, double v) {
  if (!c || id < 0 || id >= MH_NSCALARS) return fail(c, MH_E_ARG, "bad scalar id");
  c->scal[id] = v;
  c->scal_set[id] = true;
  drop_graphs(c);  // captured launches hold the scalars by value
  return MH_OK;
}

double mh_get_scalar(const mh_ctx *c, int id) {
  if (!c || id < 0 || id >= MH_NSCALARS) return NAN;
  return c->scal[id];
}

// Every reference scalar must be set; VISLAM, FTCOND and the library
// controls are optional, and FLOWIN and RFLOW are needed only with
// IN_FLOW != 0.  Then the options this library implements.
static int check_inputs(mh_ctx *c) {
  const bool forcing = c->scal_set[MH_S_IN_FLOW] && si(c, MH_S_IN_FLOW) != 0;
  const bool repeating = c->scal_set[MH_S_IF_REPEAT] && si(c, MH_S_IF_REPEAT) != 0;
  for (int i = 0; i < MH_NSCALARS; i++)
    if (!c->scal_set[i] && i != MH_S_EXACT_ORDER && i != MH_S_USE_GRAPHS && i != MH_S_LANES && i != MH_S_MIXED &&
        i != MH_S_FLOW_CONV && i != MH_S_TUNE && i != MH_S_SCAT_TK && i != MH_S_SMOOTH_WAVES &&
        i != MH_S_VISLAM &&
        i != MH_S_FTCOND &&
        (forcing || (i != MH_S_FLOWIN && i != MH_S_RFLOW)) &&
        (repeating || (i != MH_S_NINMOD && i != MH_S_RFINBC && i != MH_S_PO_IN_MID && i != MH_S_VT_IN_MID)))
      return fail(c, MH_E_STATE, std::string("scalar not set: ") + kScalarNames[i]);
  const Params p = params_of(c);
  if (p.IFGAS != 0) return fail(c, MH_E_UNSUPPORTED, "real-gas properties (IFGAS=1)");
  if (p.ITIMST != 3) return fail(c, MH_E_UNSUPPORTED, "ITIMST other than 3/-3 (artificial compressibility)");
  if (p.IF_REPEAT != 0 && p.NINMOD < 1) return fail(c, MH_E_ARG, "IF_REPEAT set with NINMOD < 1");
  if (p.ILOS != 0 && p.ILOS != 10 && p.ILOS < 100)
    return fail(c, MH_E_UNSUPPORTED, "ILOS other than 0, 10 (LOSS), 100-199 (NEW_LOSS) or >= 200 (SPAL_LOSS)");
  if (p.ILOS != 0 && p.NLOS < 1) return fail(c, MH_E_ARG, "NLOS < 1");
  if (p.IMID < 1 || p.IMID > c->g.IM || p.KMID < 1 || p.KMID > c->g.KM) return fail(c, MH_E_ARG, "IMID/KMID");
  c->g.IMID = p.IMID;
  c->g.KMID = p.KMID;
  return MH_OK;
}

int mh_loss_init(mh_ctx *c) {
  if (!c) return MH_E_ARG;
  int r = check_inputs(c);
  if (r) return r;
  const Params p = params_of(c);
  if (p.ILOS == 0) return MH_OK;
  float vf[2] = {(float)c->scal[MH_S_VISLAM], (float)c->scal[MH_S_FTCOND]};
  if ((r = h2d(c, c->d.SC + SC_VISLAM, vf, sizeof vf))) return r;
  const float one = 1.0f;  // SETUP's call has FMIXUP = 1 (mf:7350)
  if ((r = h2d(c, c->d.SC + SC_FMIXUP, &one, sizeof one))) return r;
  if (p.ILOS == 10) {
    if ((r = loss_tables(c)) || (r = loss(c, c->d, 1.0f, true))) return r;
  } else {
    if ((r = sa_maps(c)) || (r = viscous(c, c->d, p.ILOS, 1.0f, true, true))) return r;
  }
  if ((r = d2h(c, vf, c->d.SC + SC_VISLAM, sizeof vf))) return r;
  c->scal[MH_S_VISLAM] = vf[0];
  c->scal[MH_S_FTCOND] = vf[1];
  c->scal_set[MH_S_VISLAM] = c->scal_set[MH_S_FTCOND] = true;
  return MH_OK;
}

// SET_XLENGTH's launches on work buffers the caller provides: th (the nodal
// THETA, n floats), dist (the nodal wall distance, n), xl4 (the four per-row
// limits, 4 NROWS) and xlim (the per-J limit, JM).  Every kernel writes each
// work value before it reads it, so the buffers need no clearing.
static int xlength_run(mh_ctx *c, float *th, float *dist, float *xl4, float *xlim, const float *theta,
                       const float *xllim_in, const float *xllim_le, const float *xllim_te, const float *xllim_dn) {
  const Params p = params_of(c);
  const Geo g = c->g;
  const size_t n = (size_t)g.IM * g.JM * g.KM;
  const int nr = g.NROWS;
  const hipStream_t st = c->stream;
  HIPCHK(hipMemcpyAsync(th, theta, n * 4, hipMemcpyHostToDevice, st));
  const float *rows[4] = {xllim_in, xllim_le, xllim_te, xllim_dn};
  for (int q = 0; q < 4; q++)
    HIPCHK(hipMemcpyAsync(xl4 + (size_t)q * nr, rows[q], (size_t)nr * 4, hipMemcpyHostToDevice, st));
  const Dev &d = c->d;
  const unsigned nb = (unsigned)((n + 255) / 256);
  LAUNCH(k_xl_dist, dim3(nb), dim3(256), 0, st, g, d, th, dist, p.IBOUND);
  LCHK("xl_dist");
  LAUNCH(k_xl_limit, dim3((g.JM + 63) / 64), dim3(64), 0, st, g, d, dist, xl4, xlim);
  LCHK("xl_limit");
  const float dblend = 0.6666666f, expon = dblend / (1.0f - dblend), fblend = (1.0f - dblend) * powf(dblend, expon);
  LAUNCH(k_xl_cells, dim3(nb), dim3(256), 0, st, g, d, dist, xlim, dblend, expon, fblend);
  LCHK("xl_cells");
  HIPCHK(hipStreamSynchronize(st));
  return MH_OK;
}

// SET_XLENGTH on the device.  The work buffers live for this call only.
int mh_set_xlength(mh_ctx *c, const float *theta, const float *xllim_in, const float *xllim_le,
                   const float *xllim_te, const float *xllim_dn) {
  if (!c) return MH_E_ARG;
  if (!theta || !xllim_in || !xllim_le || !xllim_te || !xllim_dn) return fail(c, MH_E_ARG, "null argument");
  int r = check_inputs(c);
  if (r) return r;
  const Geo g = c->g;
  const size_t n = (size_t)g.IM * g.JM * g.KM;
  struct Bufs {
    float *th = nullptr, *dist = nullptr, *xl4 = nullptr, *xlim = nullptr;
    ~Bufs() {
      (void)hipFree(th);
      (void)hipFree(dist);
      (void)hipFree(xl4);
      (void)hipFree(xlim);
    }
  } b;
  HIPCHK(hipMalloc(&b.th, n * 4));
  HIPCHK(hipMalloc(&b.dist, n * 4));
  HIPCHK(hipMalloc(&b.xl4, 4 * (size_t)g.NROWS * 4));
  HIPCHK(hipMalloc(&b.xlim, (size_t)g.JM * 4));
  return xlength_run(c, b.th, b.dist, b.xl4, b.xlim, theta, xllim_in, xllim_le, xllim_te, xllim_dn);
}

// The same on this device's file-scope scratch array: no allocation.  The
// array is shared by every context on the device, so overlapping calls on
// one GPU race (the caller serialises them).
int mh_set_xlength_noalloc(mh_ctx *c, const float *theta, const float *xllim_in, const float *xllim_le,
                           const float *xllim_te, const float *xllim_dn) {
  if (!c) return MH_E_ARG;
  if (!theta || !xllim_in || !xllim_le || !xllim_te || !xllim_dn) return fail(c, MH_E_ARG, "null argument");
  const Geo g = c->g;
  const size_t n = (size_t)g.IM * g.JM * g.KM;
  if (n > (size_t)MH_XLENGTH_SCRATCH_NODES || g.JM > MH_XLENGTH_SCRATCH_J)
    return fail(c, MH_E_ARG,
                "mh_set_xlength_noalloc: " + std::to_string(n) + " nodes, JM = " + std::to_string(g.JM) +
                    ": the scratch array holds at most " + std::to_string(MH_XLENGTH_SCRATCH_NODES) +
                    " nodes and JM <= " + std::to_string(MH_XLENGTH_SCRATCH_J) + " (use mh_set_xlength)");
  if (!c->xl_scratch) return fail(c, MH_E_STATE, "mh_set_xlength_noalloc: no scratch address for this device");
  int r = check_inputs(c);
  if (r) return r;
  float *th = c->xl_scratch, *dist = th + MH_XLENGTH_SCRATCH_NODES;
  float *xl4 = dist + MH_XLENGTH_SCRATCH_NODES, *xlim = xl4 + kXlRowSlots;
  return xlength_run(c, th, dist, xl4, xlim, theta, xllim_in, xllim_le, xllim_te, xllim_dn);
}

// Rows with IF_CUSP = 2: their separation force runs each step; every node
// and J it reads must lie in the grid (I = 1..6 and IM-6..IM; J from one
// before the start points to N_WAKE past the trailing edge).
static int check_te_sep(mh_ctx *c) {
  const Geo &g = c->g;
  const int NR = g.NROWS;
  std::vector<int> ifc(NR), nu1(NR), num(NR), nwk(NR), jte(NR + 1);
  int r;
  if ((r = d2h(c, ifc.data(), c->d.IF_CUSP, NR * 4)) || (r = d2h(c, nu1.data(), c->d.NUP_I1, NR * 4)) ||
      (r = d2h(c, num.data(), c->d.NUP_IM, NR * 4)) || (r = d2h(c, nwk.data(), c->d.N_WAKE, NR * 4)) ||
      (r = d2h(c, jte.data(), c->d.JTE, (NR + 1) * 4)))
    return r;
  c->te_sep = false;
  for (int n = 0; n < NR; n++) {
    if (ifc[n] != 2) continue;
    const int j1 = jte[n] - nu1[n], j2 = jte[n] - num[n], jend = jte[n] + nwk[n];
    if (g.IM < 7 || std::min(j1, j2) < 2 || std::max(j1, j2) + 1 > g.JM || jend > g.JM || nwk[n] < 0)
      return fail(c, MH_E_ARG, "IF_CUSP = 2 row " + std::to_string(n + 1) +
                                   ": NUP_I1, NUP_IM or N_WAKE reach outside the grid");
    c->te_sep = true;
  }
  return MH_OK;
}

int mh_start(mh_ctx *c) {
  if (!c) return MH_E_ARG;
  int r = check_inputs(c);
  if (r) return r;
  const Params p = params_of(c);
  if ((r = build_lists(c))) return r;
  if ((r = check_te_sep(c))) return r;
  if (p.ILOS != 0) {
    if (p.ILOS == 10 && (r = loss_tables(c))) return r;
    if (p.ILOS >= 100 && (r = sa_maps(c))) return r;
    float vf[2] = {(float)c->scal[MH_S_VISLAM], (float)c->scal[MH_S_FTCOND]};
    if ((r = h2d(c, c->d.SC + SC_VISLAM, vf, sizeof vf))) return r;
  }
  // LOOP prologue, mf:3674-3714
  const Geo &g = c->g;
  std::vector<float> po1(g.KM), p1(g.KM);
  float pa = 0, pb = 0, ra = 0, rb = 0;
  const size_t nA = (size_t)(p.IMID - 1) + (size_t)g.IM * (size_t)g.JM * (p.KMID - 1);
  const size_t nB = nA + (size_t)g.IM * (g.JM - 1);
  if ((r = d2h(c, po1.data(), c->d.PO1, g.KM * 4))) return r;
  for (int K = 1; K <= g.KM; K++)
    if ((r = d2h(c, &p1[K - 1], c->d.P + (size_t)g.IM * g.JM * (K - 1), 4))) return r;
  if ((r = d2h(c, &pa, c->d.P + nA, 4)) || (r = d2h(c, &pb, c->d.P + nB, 4)) || (r = d2h(c, &ra, c->d.RO + nA, 4)) ||
      (r = d2h(c, &rb, c->d.RO + nB, 4)))
    return r;
  c->ROMIN = fminf(ra, rb);
  c->ROLIM = 0.25f * c->ROMIN;
  std::vector<float> plim(g.KM);
  for (int K = 1; K <= g.KM; K++) plim[K - 1] = 0.9f * po1[K - 1] + 0.1f * p1[K - 1];
  if ((r = h2d(c, c->d.SC + SC_PLIM, plim.data(), g.KM * 4))) return r;
  c->PREFF = 0.5f * (pa + pb);
  c->SFXIN_s = sf(c, MH_S_SFXIN) * sf(c, MH_S_CFL) / 0.5f;
  c->SFTIN_s = sf(c, MH_S_SFTIN) * sf(c, MH_S_CFL) / 0.5f;
  // the mixed form: the SSS histories kept in BF16 from here on
  const bool mixed = c->scal[MH_S_MIXED] != 0.0;
  if (mixed && p.EXACT) return fail(c, MH_E_UNSUPPORTED, "the mixed form (MIXED = 1) needs EXACT_ORDER = 0");
  if (c->mixed && !mixed)  // back to FP32 storage: take the histories as they are
    for (int i = 0; i < kHist; i++)
      if ((r = hist_convert(c, i, false))) return r;
  c->mixed = false;
  if (mixed) {
    const size_t n = (size_t)g.IM * g.JM * g.KM;
    for (int i = 0; i < kHist; i++) {
      if (!c->h16[i] && (r = dalloc(c, &c->h16[i], n))) return r;
      if ((r = hist_convert(c, i, true))) return r;
    }
    // STORE's BF16 copies (every call writes them before they are read)
    if (!c->store16 && (r = dalloc(c, &c->store16, n))) return r;
    c->mixed = true;
  }
  if (!p.EXACT && (r = make_lanes(c))) return r;
  if (c->mixed)
    for (int l = 0; l < c->nlanes; l++)
      if (!c->lanes[l].STORE16 && (r = dalloc(c, &c->lanes[l].STORE16, (size_t)g.IM * g.JM * g.KM))) return r;
  if ((r = autotune(c))) return r;
  c->ifend = 0;
  c->nstep = 0;
  c->hstep = 0;
  c->DAMP = 0.0f;
  c->stop_req = false;
  drop_graphs(c);  // they hold this run's scalars
  {
    // the device's step state: NSTEP = 0, the stop tests' figures as
    // before any check (mf:3697-3699), the ramps not yet started
    const std::vector<int> ctl(CTL_COUNT, 0);
    float sc[SC_NHEAD] = {};
    if ((r = d2h(c, sc, c->d.SC, sizeof sc))) return r;
    sc[SC_EMAX] = 1.0e06f;
    sc[SC_EAVG] = 1.0e06f;
    sc[SC_ECONT] = 1.f;
    sc[SC_SFT] = sc[SC_SFX] = sc[SC_DAMP] = 0.0f;
    sc[SC_FMIXUP] = 1.0f;
    sc[SC_FLOWPREV] = 0.0f;
    sc[SC_DFLOW] = 1.0e6f;
    if ((r = h2d(c, c->d.CTL, ctl.data(), CTL_COUNT * sizeof(int))) || (r = h2d(c, c->d.SC, sc, sizeof sc))) return r;
  }
  if ((r = d2h(c, c->nout, c->d.NOUT, sizeof c->nout))) return r;
  c->last = mh_resid{};
  c->started = true;
  return MH_OK;
}

// Steps are queued in chunks that end on the every-5-step checks, each chunk
// one graph launch.  The host runs at most kAhead chunks ahead of the device
// (it waits on the event of the chunk kAhead back), which bounds the queue
// depth; while queuing, the device's end flag is read back without blocking
// (a copy into pinned memory, polled with an event), so queuing stops soon
// after a run ends by convergence; the steps queued beyond it are not taken
// on the device.
int mh_step(mh_ctx *c, int nsteps, mh_resid *out) {
  if (!c) return MH_E_ARG;
  if (!c->started) return fail(c, MH_E_STATE, "mh_step before mh_start");
  if (nsteps < 0) return fail(c, MH_E_ARG, "nsteps < 0");
  const Params p = params_of(c);
  int r;
  bool end = c->ifend != 0, polling = false;
  int queued = 0;
  std::vector<StepPlan> plan;
  while (queued < nsteps && !end) {
    const int hn = c->hstep;
    const int len = std::min(nsteps - queued, 5 - hn % 5);
    plan.clear();
    for (int i = 1; i <= len; i++) {
      const int n = hn + i;
      plan.push_back(plan_step(c, n));
      if (n == p.NSTEPS_MAX || (c->stop_req && n % 10 == 0)) {
        end = true;
        break;
      }
    }
    hipEvent_t &slot = c->ev_ahead[c->nchunk % kAhead];
    if (c->nchunk >= kAhead) HIPCHK(hipEventSynchronize(slot));
    if ((r = enqueue_chunk(c, plan))) return r;
    HIPCHK(hipEventRecord(slot, c->stream));
    c->nchunk++;
    c->hstep = hn + (int)plan.size();
    queued += (int)plan.size();
    if (!polling) {
      HIPCHK(hipMemcpyAsync(c->ctl_host, c->d.CTL, CTL_COUNT * sizeof(int), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipEventRecord(c->ev_poll, c->stream));
      polling = true;
    } else if (hipEventQuery(c->ev_poll) == hipSuccess) {
      polling = false;
      if (c->ctl_host[CTL_DONE] || c->ctl_host[CTL_IFEND]) end = true;
    }
    if (c->prof_on && c->prof_pending.size() > 4096 && (r = prof_collect(c))) return r;
  }
  r = read_back(c);
  c->nchunk = 0;        // the stream is drained: no chunk is outstanding
  c->hstep = c->nstep;  // the device may have ended the run before the host's last queued step
  if (out) *out = c->last;
  return r;
}

int mh_request_stop(mh_ctx *c) {
  if (!c) return MH_E_ARG;
  if (!c->started) return fail(c, MH_E_STATE, "mh_request_stop before mh_start");
  c->stop_req = true;
  const int one = 1;
  return h2d(c, c->d.CTL + CTL_STOPREQ, &one, sizeof one);
}

int mh_step_batch(mh_ctx *const *ctxs, int n, int nsteps, mh_resid *out, int *failed) {
  if (!ctxs || !out || n < 1 || nsteps < 0) return MH_E_ARG;
  for (int i = 0; i < n; i++) {
    if (!ctxs[i]) return MH_E_ARG;
    for (int k = 0; k < i; k++)
      if (ctxs[k] == ctxs[i]) return fail(ctxs[i], MH_E_ARG, "mh_step_batch: a context appears twice");
  }
  // whole contexts overlap here, each on its own stream
  std::vector<int> rc(n, MH_OK);
  std::vector<std::thread> th;
  th.reserve(n);
  for (int i = 0; i < n; i++) th.emplace_back([&, i] { rc[i] = mh_step(ctxs[i], nsteps, &out[i]); });
  for (auto &t : th) t.join();
  for (int i = 0; i < n; i++)
    if (rc[i] != MH_OK) {
      if (failed) *failed = i;
      return rc[i];
    }
  return MH_OK;
}

int mh_integrals(mh_ctx *c, double *out, long long n) {
  if (!c || !out || n != 7LL * c->g.JM) return fail(c, MH_E_ARG, "integrals: need 7*JM doubles");
  const Params p = params_of(c);
  LAUNCH(k_integrals, dim3(c->g.JM), dim3(256), 0, c->stream, c->g, c->d, p);
  LCHK("integrals");
  HIPCHK(hipMemcpyAsync(out, c->d.INTEG, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MH_OK;
}

// The stages of mf:3600-3614 from NSTAGE (the reference's default when it
// was never set), each stage's planes as EFICOOL takes them (mf:11271-11295),
// then k_integrals and k_eficool.
int mh_eficool(mh_ctx *c, mh_perf *out, int cap, int *n) {
  if (!c || !n || cap < 0 || (cap > 0 && !out)) return fail(c, MH_E_ARG, "mh_eficool: bad arguments");
  const Geo &g = c->g;
  const int NR = g.NROWS, JM = g.JM;
  std::vector<int> nst(NR), jstart(NR + 1), jmix(NR + 1), nrow(JM);
  int r;
  if ((r = d2h(c, nst.data(), c->d.NSTAGE, NR * 4)) || (r = d2h(c, jstart.data(), c->d.JSTART, (NR + 1) * 4)) ||
      (r = d2h(c, jmix.data(), c->d.JMIX, (NR + 1) * 4)) || (r = d2h(c, nrow.data(), c->d.NROW, JM * 4)))
    return r;
  if (std::all_of(nst.begin(), nst.end(), [](int v) { return v == 0; }))
    for (int N = 1; N <= NR; N++) nst[N - 1] = 1 + (N - 1) / 2;
  const int NSTAGES = nst[NR - 1];
  if (NSTAGES < 1 || NSTAGES > NR) return fail(c, MH_E_ARG, "mh_eficool: NSTAGE of the last row out of 1..NROWS");
  std::vector<int> js(NSTAGES + 2, 0), je(NSTAGES + 2, 0);
  js[1] = 1;
  for (int N = 1; N <= NR; N++) {
    const int a = nst[N - 1], b = N < NR ? nst[N] : NSTAGES + 1;
    if (a < 1 || a > NSTAGES || b < 1 || b > NSTAGES + 1) return fail(c, MH_E_ARG, "mh_eficool: NSTAGE out of range");
    if (a != b) {
      je[a] = jmix[N - 1];
      js[b] = N < NR ? jstart[N] : JM + 1;
    }
  }
  std::vector<int> tab;
  auto add = [&](int j1, int j2m1, int jm, int stage) {
    tab.insert(tab.end(), {j1, j2m1, jm, stage});
  };
  for (int S = 1; S <= NSTAGES; S++) {
    const int J1 = js[S];
    const int J2M1 = S == NSTAGES ? JM - 1 : je[S] + 1;
    if (J1 < 1 || J1 + 1 > JM) return fail(c, MH_E_ARG, "mh_eficool: a stage's first plane is out of the grid");
    const int NRW = nrow[J1];  // NROW(J1+1)
    if (NRW < 1 || NRW > NR) return fail(c, MH_E_ARG, "mh_eficool: NROW");
    add(J1, J2M1, jmix[NRW - 1], S);
  }
  if (NSTAGES > 1) add(2, JM - 1, (2 + JM) / 2, 0);
  const int nrec = (int)tab.size() / 4;
  for (int i = 0; i < nrec; i++)
    for (int q = 0; q < 3; q++)
      if (tab[4 * i + q] < 1 || tab[4 * i + q] > JM) return fail(c, MH_E_ARG, "mh_eficool: a plane is out of the grid");
  *n = nrec;
  if (cap < nrec) return fail(c, MH_E_ARG, "mh_eficool: room for " + std::to_string(nrec) + " records needed");
  if (!c->eff_tab && ((r = dalloc(c, &c->eff_tab, 4 * (size_t)(NR + 1))) ||
                      (r = dalloc(c, &c->eff_out, (size_t)kEffFields * (NR + 1)))))
    return r;
  if ((r = h2d(c, c->eff_tab, tab.data(), tab.size() * 4))) return r;
  const Params p = params_of(c);
  LAUNCH(k_integrals, dim3(JM), dim3(256), 0, c->stream, c->g, c->d, p);
  LCHK("integrals");
  LAUNCH(k_eficool, dim3(1), dim3(64), 0, c->stream, c->g, c->d, p, c->eff_tab, nrec, c->eff_out);
  LCHK("eficool");
  std::vector<double> v((size_t)kEffFields * nrec);
  if ((r = d2h(c, v.data(), c->eff_out, v.size() * 8))) return r;
  for (int i = 0; i < nrec; i++) {
    const double *f = &v[(size_t)kEffFields * i];
    mh_perf &o = out[i];
    o = mh_perf{};
    o.stage = tab[4 * i + 3];
    o.j1 = tab[4 * i];
    o.j2m1 = tab[4 * i + 1];
    o.jmix = tab[4 * i + 2];
    o.compressor = f[17] != 0.0;
    o.po_in = f[0]; o.po_out = f[1]; o.ps_in = f[2]; o.ps_out = f[3]; o.to_in = f[4]; o.to_out = f[5];
    o.flow_in = f[6]; o.flow_out = f[7]; o.po_ratio = f[8]; o.eta_tt = f[9]; o.eta_ts = f[10];
    o.eta_poly = f[11]; o.reaction = f[12]; o.reaction_hub = f[13]; o.reaction_mid = f[14];
    o.reaction_tip = f[15]; o.work_kw = f[16];
  }
  return MH_OK;
}

int mh_mix_bconds(mh_ctx *c, double *out, long long n) {
  if (!c || !out || n != (long long)(c->g.NROWS + 1) * 8 * (c->g.KM + 1))
    return fail(c, MH_E_ARG, "mh_mix_bconds: need (NROWS+1)*8*(KM+1) doubles");
  const Params p = params_of(c);
  LAUNCH(k_mix_bconds, dim3((c->g.NROWS + 1) * 8), dim3(64), (c->g.KM + 1) * sizeof(double), c->stream, c->g, c->d,
         p);
  LCHK("mix_bconds");
  HIPCHK(hipMemcpyAsync(out, c->d.MIXB, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MH_OK;
}

int mh_profile_enable(mh_ctx *c, int on) {
  if (!c) return MH_E_ARG;
  int r = prof_collect(c);
  if (r) return r;
  c->prof_on = on != 0;
  c->prof_tot.clear();
  return MH_OK;
}

int mh_profile_report(mh_ctx *c, char *buf, int cap) {
  if (!c || !buf || cap <= 0) return MH_E_ARG;
  int r = prof_collect(c);
  if (r) return r;
  std::string out;
  for (auto &t : c->prof_tot) {
    char line[160];
    snprintf(line, sizeof line, "%s %lld %.6f\n", t.first.c_str(), t.second.first, t.second.second);
    out += line;
  }
  int n = (int)std::min<size_t>(out.size(), (size_t)cap - 1);
  memcpy(buf, out.data(), n);
  buf[n] = 0;
  return n;
}

int mh_sync(mh_ctx *c) {
  if (!c) return MH_E_ARG;
  HIPCHK(hipStreamSynchronize(c->stream));
  return MH_OK;
}

}  // extern "C"
