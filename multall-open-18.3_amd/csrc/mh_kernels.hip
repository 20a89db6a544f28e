// HIP kernels for one time step of the reference's LOOP (gfx950).
//
// Every kernel names the reference lines it implements
// (Multall-open-18.3/multall-18.3/multall-open-18.3.f, "mf:").  The
// statement-level arithmetic follows the reference's FP32 expression order
// (the library is built with -ffp-contract=off), so pointwise and stencil
// phases reproduce the reference bit-for-bit; reductions either follow the
// reference's sequential order (EXACT mode) or use deterministic FP64 trees.
//
// Index helpers take the reference's 1-based (I,J,K).  Node-range kernels
// use blocks of 64 (I) x 4 (J) threads and one grid z-slice per K, so a
// wavefront covers one pitchwise line: coalesced 256-byte rows of the
// I-fastest layout.
#include "mh_internal.h"

namespace mh {

#define IX(I, J, K) ((size_t)((I) - 1) + (size_t)g.IM * ((size_t)((J) - 1) + (size_t)g.JM * (size_t)((K) - 1)))
#define JKX(J, K) ((size_t)((J) - 1) + (size_t)g.JM * (size_t)((K) - 1))
#define IKX(I, K) ((size_t)((I) - 1) + (size_t)g.IM * (size_t)((K) - 1))
#define IJX(I, J) ((size_t)((I) - 1) + (size_t)g.IM * (size_t)((J) - 1))
#define NODE_IJK                                           \
  const int I = blockIdx.x * blockDim.x + threadIdx.x + 1; \
  const int J = blockIdx.y * blockDim.y + threadIdx.y + 1; \
  const int K = blockIdx.z + 1;                            \
  if (I > g.IM || J > g.JM) return;


// Left-to-right FP32 sum of x[1..n] (n <= 256) on one wavefront, in the
// reference's sequential order: lane l holds x[l+1+64c] in v[c].  Every lane
// returns the same value.  The shuffles are independent; only the adds chain.
__device__ __forceinline__ float wave_ordered_sum(const float (&v)[4], int n) {
  float s = 0.0f;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int m = min(64, n - 64 * c);
    for (int l = 0; l < m; l++) s = s + __shfl(v[c], l, 64);
  }
  return s;
}

// Sum of x[1..n] (n <= 256) held as in wave_ordered_sum.  EXACT mode keeps
// the reference's order; otherwise a deterministic FP64 butterfly (every
// lane ends with the same bits: each level adds the same two operands).
__device__ __forceinline__ float line_sum(const float (&v)[4], int n, int exact) {
  if (exact) return wave_ordered_sum(v, n);
  double s = ((double)v[0] + (double)v[1]) + ((double)v[2] + (double)v[3]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  return (float)s;
}

// The wavefront's index in its workgroup, as a scalar (SGPR) value, so that
// every line index derived from it is computed once per wavefront on the
// scalar unit rather than per lane.
// FP64 sum over the 64 lanes of a wavefront (every lane gets it)
__device__ __forceinline__ double wave_sum_d(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

// FP32 transcendentals as the FP64 function rounded once to FP32: the
// correctly rounded result except at rounding ties, which is what the host
// libm's FP32 functions (the reference's) return for all but rare inputs.
// A 1-ulp difference can be amplified downstream: the inlet temperature's
// difference from TO1 gives the inlet speed, for one.
__device__ __forceinline__ float pow_r(float x, float y) { return (float)pow((double)x, (double)y); }
__device__ __forceinline__ float log_r(float x) { return (float)log((double)x); }
__device__ __forceinline__ float sin_r(float x) { return (float)sin((double)x); }
__device__ __forceinline__ float cos_r(float x) { return (float)cos((double)x); }
__device__ __forceinline__ float asin_r(float x) { return (float)asin((double)x); }
__device__ __forceinline__ float atan2_r(float y, float x) { return (float)atan2((double)y, (double)x); }
__device__ __forceinline__ float tanh_r(float x) { return (float)tanh((double)x); }

// XCD-aware workgroup order.  Workgroups are dispatched round-robin over the
// 8 XCDs, each with its own L2; this bijection on [0, n) gives XCD x the
// contiguous range of logical ids it would otherwise share with the other
// seven, so neighbouring J planes (stencil halos) meet in one L2.
__device__ __forceinline__ int xcd_block(int b, int n) {
  constexpr int kXcd = 8;
  const int q = n / kXcd, r = n % kXcd, x = b % kXcd, l = b / kXcd;
  return x < r ? x * (q + 1) + l : r * (q + 1) + (x - r) * q + l;
}

// Batched launches: a workgroup works on variable e.q[z] (gridDim.z
// variables).  The XCD-aware order runs over the whole grid with the
// variable fastest, so one XCD takes a contiguous range of tiles for all
// variables at once and the arrays they share (the time steps, the grid)
// are fetched into its L2 once.  Returns the logical tile; z is
// wavefront-uniform.
__device__ __forceinline__ int eq_block(int &z) {
  const int nz = gridDim.z;
  const int b = xcd_block(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), gridDim.x * gridDim.y * nz);
  const int t = b / nz;
  z = b - t * nz;
  return t;
}
// e.q[z] by uniform selects (a dynamic index would copy the set to scratch)
__device__ __forceinline__ const EqLane &eq_pick(const EqSet &e, int z) {
  return z == 0 ? e.q[0] : z == 1 ? e.q[1] : e.q[2];
}
// the context's arrays with variable q's flux and TSTEP scratch
__device__ __forceinline__ Dev eq_dev(const Dev &d0, const EqLane &q) {
  Dev d = d0;
  d.XFLUX = q.XFLUX;
  d.TFLUX = q.TFLUX;
  d.RFLUX = q.RFLUX;
  d.SOURCE = q.SOURCE;
  d.STORE = q.STORE;
  d.DSC = q.DSC;
  d.TMP = q.TMP;
  d.B1CHG = q.B1CHG;
  d.B2CHG = q.B2CHG;
  d.SBCHG = q.SBCHG;
  d.AVG = q.AVG;
  d.RED = q.RED;
  return d;
}

// ---------------------------------------------------------------------------
// primitive variables, mf:3826-3835 and mf:3871-3897 (ideal gas)

// (node kernels check the step's gate after their loads, before their
// first store, so the loads do not wait on it)
__global__ void k_prim(Geo g, Dev d, Params p) {
  NODE_IJK
  if (J < 2) return;
  const size_t n = IX(I, J, K);
  const float roe = d.ROE[n];
  float RRNOW = 1 / d.R[JKX(J, K)];
  float RONOW = d.RO[n];
  float RECIP = 1.0f / RONOW;
  float vx = d.ROVX[n] * RECIP;
  float vr = d.ROVR[n] * RECIP;
  float rovt = d.RORVT[n] * RRNOW;
  float vt = rovt * RECIP;
  if (gated(d)) return;
  d.VX[n] = vx;
  d.VR[n] = vr;
  d.ROVT[n] = rovt;
  d.VT[n] = vt;
  if (J <= g.JMM1) {
    float EKE = .5f * (vr * vr + vt * vt + vx * vx);
    float TSTATIC = (roe / RONOW - EKE) / p.CV;
    d.P[n] = RONOW * p.RGAS * TSTATIC;
    d.HO[n] = p.CP * TSTATIC + EKE;
    d.T_STATIC[n] = TSTATIC;
  }
}

// ---------------------------------------------------------------------------
// mass fluxes, relative swirl, periodic and blade-surface flux conditions,
// mf:3960-4190 (no bleed, cooling or shroud terms)

__device__ __forceinline__ float rowt_at(const Geo &g, const Dev &d, int I, int J, int K) {
  const size_t n = IX(I, J, K);
  return d.ROVT[n] - d.UBLADE[JKX(J, K)] * d.RO[n];
}

__device__ __forceinline__ float flowt_raw(const Geo &g, const Dev &d, int I, int J, int K) {
  float AVGRVX = d.ROVX[IX(I, J, K)] + d.ROVX[IX(I, J - 1, K)] + d.ROVX[IX(I, J - 1, K + 1)] + d.ROVX[IX(I, J, K + 1)];
  float AVGRVT = rowt_at(g, d, I, J, K) + rowt_at(g, d, I, J - 1, K) + rowt_at(g, d, I, J - 1, K + 1) +
                 rowt_at(g, d, I, J, K + 1);
  float AVGRVR = d.ROVR[IX(I, J, K)] + d.ROVR[IX(I, J - 1, K)] + d.ROVR[IX(I, J - 1, K + 1)] + d.ROVR[IX(I, J, K + 1)];
  return (AVGRVX * d.ABX[IX(I, J, K)] + AVGRVT * d.ABT[JKX(J, K)] + AVGRVR * d.ABR[IX(I, J, K)]) * 0.25f;
}

// The node's mass fluxes (zero on the faces not formed there) and relative
// swirl, formed without storing.
struct MassFlux {
  float fx = 0.f, ft = 0.f, fr = 0.f, rowt = 0.f, wt = 0.f;
  bool hx = false, ht = false, hr = false;
};
__device__ __forceinline__ MassFlux massflux_vals(const Geo &g, const Dev &d, int I, int J, int K) {
  MassFlux m;
  const size_t n = IX(I, J, K);
  if (I <= g.IMM1 && K <= g.KMM1) {
    float AVGRVX = d.ROVX[IX(I, J, K)] + d.ROVX[IX(I, J, K + 1)] + d.ROVX[IX(I + 1, J, K)] + d.ROVX[IX(I + 1, J, K + 1)];
    float AVGRVR = d.ROVR[IX(I, J, K)] + d.ROVR[IX(I, J, K + 1)] + d.ROVR[IX(I + 1, J, K)] + d.ROVR[IX(I + 1, J, K + 1)];
    m.fx = 0.25f * (AVGRVX * d.AQX[n] + AVGRVR * d.AQR[n]);
    m.hx = true;
  }
  m.rowt = rowt_at(g, d, I, J, K);
  m.wt = m.rowt / d.RO[n];
  if (J >= 2 && K <= g.KMM1) {
    if (d.IND[J - 1] != 1 && (I == 1 || I == g.IM)) {
      m.ft = 0.5f * (flowt_raw(g, d, 1, J, K) + flowt_raw(g, d, g.IM, J, K));
    } else {
      m.ft = flowt_raw(g, d, I, J, K);
      if (d.IND[J - 1] == 1 && (I == 1 || I == g.IM)) {
        int NR = d.NROW[J - 1];
        if (K >= d.KS1[NR - 1] && K <= d.KS2[NR - 1]) m.ft = 0.0f;
      }
    }
    m.ht = true;
  }
  if (J >= 2 && I <= g.IMM1) {
    if (K >= 2 && K <= g.KMM1) {
      float AVGRVX = d.ROVX[IX(I, J, K)] + d.ROVX[IX(I, J - 1, K)] + d.ROVX[IX(I + 1, J - 1, K)] + d.ROVX[IX(I + 1, J, K)];
      float AVGRVR = d.ROVR[IX(I, J, K)] + d.ROVR[IX(I, J - 1, K)] + d.ROVR[IX(I + 1, J - 1, K)] + d.ROVR[IX(I + 1, J, K)];
      m.fr = 0.25f * (AVGRVX * d.ASX[n] + AVGRVR * d.ASR[n]);
    }
    m.hr = true;
  }
  return m;
}
// FLOWX, FLOWT, FLOWR, ROWT, WT and, with dup, continuity's fluxes (the
// mass fluxes, no source) in d's flux arrays
__device__ __forceinline__ void massflux_store(const Geo &g, const Dev &d, const MassFlux &m, size_t n, bool dup) {
  if (m.hx) {
    d.FLOWX[n] = m.fx;
    if (dup) {
      d.SOURCE[n] = 0.0f;
      d.XFLUX[n] = m.fx;
    }
  }
  d.ROWT[n] = m.rowt;
  d.WT[n] = m.wt;
  if (m.ht) {
    d.FLOWT[n] = m.ft;
    if (dup) d.TFLUX[n] = m.ft;
  }
  if (m.hr) {
    d.FLOWR[n] = m.fr;
    if (dup) d.RFLUX[n] = m.fr;
  }
}
__global__ void k_massflux(Geo g, Dev d) {
  if (gated(d)) return;
  NODE_IJK
  massflux_store(g, d, massflux_vals(g, d, I, J, K), IX(I, J, K), true);
}

// ---------------------------------------------------------------------------
// energy fluxes, mf:4357-4456

// FLT: FLOWT(I,J,K)
__device__ __forceinline__ float tflux_energy(const Geo &g, const Dev &d, int I, int J, int K, float FLT) {
  float AVGHO = d.HO[IX(I, J, K)] + d.HO[IX(I, J - 1, K)] + d.HO[IX(I, J - 1, K + 1)] + d.HO[IX(I, J, K + 1)];
  float AVGPB = d.PEFF[IX(I, J, K)] + d.PEFF[IX(I, J - 1, K)] + d.PEFF[IX(I, J - 1, K + 1)] + d.PEFF[IX(I, J, K + 1)];
  return (FLT * AVGHO + d.WRABT[JKX(J, K)] * AVGPB) * 0.25f;
}

// The face fluxes and source a flux formula forms at one node (has_*: the
// faces it forms there), so that a thread can form several before storing.
struct FaceFlux {
  float x = 0.f, r = 0.f, t = 0.f, s = 0.f;
  bool hx = false, hr = false, ht = false, hs = false;
};
__device__ __forceinline__ void store_faces(const FaceFlux &v, float *XF, float *RF, float *TF, float *SRC, size_t n) {
  if (v.hx) XF[n] = v.x;
  if (v.hr) RF[n] = v.r;
  if (v.ht) TF[n] = v.t;
  if (v.hs) SRC[n] = v.s;
}

// fx, ft, fr: the node's FLOWX, FLOWT, FLOWR.  On a periodic line FLOWT is
// the same at I = 1 and IM (their average), so ft serves both ends.
__device__ __forceinline__ FaceFlux energy_faces(const Geo &g, const Dev &d, int I, int J, int K, float fx, float ft,
                                                 float fr) {
  const size_t n = IX(I, J, K);
  FaceFlux v;
  if (I <= g.IMM1 && K <= g.KMM1) {
    float AVGHO = d.HO[IX(I, J, K)] + d.HO[IX(I, J, K + 1)] + d.HO[IX(I + 1, J, K)] + d.HO[IX(I + 1, J, K + 1)];
    v.x = fx * AVGHO * 0.25f;
    v.hx = true;
    if (J >= 2) {
      v.s = d.QSOURCE[n];
      v.hs = true;
    }
  }
  if (J >= 2 && I <= g.IMM1) {
    float AVGHO = 0.25f * (d.HO[IX(I, J, K)] + d.HO[IX(I, J - 1, K)] + d.HO[IX(I + 1, J - 1, K)] + d.HO[IX(I + 1, J, K)]);
    v.r = fr * AVGHO;
    v.hr = true;
  }
  if (J >= 2 && K <= g.KMM1) {
    if (d.IND[J - 1] != 1 && (I == 1 || I == g.IM))
      v.t = 0.5f * (tflux_energy(g, d, 1, J, K, ft) + tflux_energy(g, d, g.IM, J, K, ft));
    else
      v.t = tflux_energy(g, d, I, J, K, ft);
    v.ht = true;
  }
  return v;
}
__device__ __forceinline__ void flux_energy_at(const Geo &g, const Dev &d, int I, int J, int K) {
  const size_t n = IX(I, J, K);
  store_faces(energy_faces(g, d, I, J, K, d.FLOWX[n], d.FLOWT[n], d.FLOWR[n]), d.XFLUX, d.RFLUX, d.TFLUX, d.SOURCE, n);
}
__global__ void k_flux_energy(Geo g, Dev d) {
  if (gated(d)) return;
  NODE_IJK
  flux_energy_at(g, d, I, J, K);
}

// ---------------------------------------------------------------------------
// axial momentum fluxes, mf:4818-4896
// PEFF at a node for the periodic momentum faces: the array k_peff wrote,
// or (the batched momentum flux kernel) formed from P as k_peff forms it
struct PeffArray {
  const float *PEFF;
  __device__ float operator()(const Geo &g, int I, int J, int K) const { return PEFF[IX(I, J, K)]; }
};
struct PeffFromP {
  const float *P;
  float PREFF, F_PDOWN;
  __device__ float operator()(const Geo &g, int I, int J, int K) const {
    const int JP1 = (J == g.JM) ? J : J + 1;
    const size_t a = IX(I, J, K);
    return P[a] - PREFF + F_PDOWN * (P[IX(I, JP1, K)] - P[a]);
  }
};


__device__ __forceinline__ float tflux_axial(const Geo &g, const Dev &d, int I, int J, int K) {
  float AVGVX = d.VX[IX(I, J, K)] + d.VX[IX(I, J - 1, K)] + d.VX[IX(I, J - 1, K + 1)] + d.VX[IX(I, J, K + 1)];
  float AVGP = d.PEFF[IX(I, J, K)] + d.PEFF[IX(I, J - 1, K)] + d.PEFF[IX(I, J - 1, K + 1)] + d.PEFF[IX(I, J, K + 1)];
  return 0.25f * (d.FLOWT[IX(I, J, K)] * AVGVX + d.ABX[IX(I, J, K)] * AVGP);
}

// the periodic line's value at I = 1 and IM off the blade
template <class PE>
__device__ __forceinline__ float tflux_axial_per(const Geo &g, const Dev &d, size_t n, int J, int K, const PE &pe) {
  const int IM = g.IM;
  float AVGPCUSP = pe(g, 1, J, K) + pe(g, 1, J - 1, K) + pe(g, 1, J, K + 1) + pe(g, 1, J - 1, K + 1) +
                   pe(g, IM, J, K) + pe(g, IM, J - 1, K) + pe(g, IM, J, K + 1) + pe(g, IM, J - 1, K + 1);
  float AVGVXCUSP = d.VX[IX(1, J, K)] + d.VX[IX(1, J - 1, K)] + d.VX[IX(1, J, K + 1)] + d.VX[IX(1, J - 1, K + 1)] +
                    d.VX[IX(IM, J, K)] + d.VX[IX(IM, J - 1, K)] + d.VX[IX(IM, J, K + 1)] + d.VX[IX(IM, J - 1, K + 1)];
  return 0.125f * (d.FLOWT[n] * AVGVXCUSP + AVGPCUSP * d.ABX[n]);
}

__device__ __forceinline__ void flux_axial_at(const Geo &g, const Dev &d, int I, int J, int K) {
  const size_t n = IX(I, J, K);
  if (I <= g.IMM1 && K <= g.KMM1) {
    d.SOURCE[n] = d.XFORCE[n];
    float AVGP = d.PEFF[IX(I, J, K)] + d.PEFF[IX(I, J, K + 1)] + d.PEFF[IX(I + 1, J, K + 1)] + d.PEFF[IX(I + 1, J, K)];
    float AVGVX = d.VX[IX(I, J, K)] + d.VX[IX(I, J, K + 1)] + d.VX[IX(I + 1, J, K + 1)] + d.VX[IX(I + 1, J, K)];
    d.XFLUX[n] = 0.25f * (AVGP * d.AQX[n] + d.FLOWX[n] * AVGVX);
  }
  if (J >= 2 && I <= g.IMM1) {
    float AVGP = d.PEFF[IX(I, J, K)] + d.PEFF[IX(I, J - 1, K)] + d.PEFF[IX(I + 1, J - 1, K)] + d.PEFF[IX(I + 1, J, K)];
    float AVGVX = d.VX[IX(I, J, K)] + d.VX[IX(I, J - 1, K)] + d.VX[IX(I + 1, J - 1, K)] + d.VX[IX(I + 1, J, K)];
    d.RFLUX[n] = 0.25f * (AVGP * d.ASX[n] + d.FLOWR[n] * AVGVX);
  }
  if (J >= 2 && K <= g.KMM1) {
    float tf;
    if (d.IND[J - 1] != 1 && (I == 1 || I == g.IM)) {
      tf = tflux_axial_per(g, d, n, J, K, PeffArray{d.PEFF});
    } else {
      tf = tflux_axial(g, d, I, J, K);
    }
    d.TFLUX[n] = tf;
  }
}
__global__ void k_flux_axial(Geo g, Dev d) {
  if (gated(d)) return;
  NODE_IJK
  flux_axial_at(g, d, I, J, K);
}

// ---------------------------------------------------------------------------
// angular momentum fluxes, mf:4913-4956

template <class PE>
__device__ __forceinline__ float tflux_angular(const Geo &g, const Dev &d, int I, int J, int K, const PE &pe) {
  float AVGR = 0.25f * d.RAVG_CELL[JKX(J, K)];
  float AVGVT = d.VT[IX(I, J, K)] + d.VT[IX(I, J - 1, K)] + d.VT[IX(I, J - 1, K + 1)] + d.VT[IX(I, J, K + 1)];
  float AVGP = pe(g, I, J, K) + pe(g, I, J - 1, K) + pe(g, I, J - 1, K + 1) + pe(g, I, J, K + 1);
  return (d.FLOWT[IX(I, J, K)] * AVGVT + d.ABT[JKX(J, K)] * AVGP) * AVGR;
}

__device__ __forceinline__ void flux_angular_at(const Geo &g, const Dev &d, int I, int J, int K) {
  const size_t n = IX(I, J, K);
  if (I <= g.IMM1 && K <= g.KMM1) {
    float AVGR = 0.125f * (d.R[JKX(J, K)] + d.R[JKX(J, K + 1)]);
    float AVGVT = d.VT[IX(I, J, K)] + d.VT[IX(I, J, K + 1)] + d.VT[IX(I + 1, J, K)] + d.VT[IX(I + 1, J, K + 1)];
    d.SOURCE[n] = d.TFORCE[n];
    d.XFLUX[n] = d.FLOWX[n] * AVGR * AVGVT;
  }
  if (J >= 2 && K <= g.KMM1) {
    float tf;
    if (d.IND[J - 1] != 1 && (I == 1 || I == g.IM))
      tf = 0.5f * (tflux_angular(g, d, 1, J, K, PeffArray{d.PEFF}) + tflux_angular(g, d, g.IM, J, K, PeffArray{d.PEFF}));
    else
      tf = tflux_angular(g, d, I, J, K, PeffArray{d.PEFF});
    d.TFLUX[n] = tf;
  }
  if (J >= 2 && I <= g.IMM1) {
    float AVGR = (d.R[JKX(J, K)] + d.R[JKX(J - 1, K)]) * 0.125f;
    float AVGVT = d.VT[IX(I, J, K)] + d.VT[IX(I, J - 1, K)] + d.VT[IX(I + 1, J - 1, K)] + d.VT[IX(I + 1, J, K)];
    d.RFLUX[n] = d.FLOWR[n] * AVGVT * AVGR;
  }
}
__global__ void k_flux_angular(Geo g, Dev d) {
  if (gated(d)) return;
  NODE_IJK
  flux_angular_at(g, d, I, J, K);
}

// ---------------------------------------------------------------------------
// radial momentum fluxes and source, mf:5003-5095

// the periodic line's value at I = 1 and IM off the blade
template <class PE>
__device__ __forceinline__ float tflux_radial_per(const Geo &g, const Dev &d, size_t n, int J, int K, const PE &pe) {
  const int IM = g.IM;
  float AVGVR1 = d.VR[IX(1, J, K)] + d.VR[IX(1, J - 1, K)] + d.VR[IX(1, J - 1, K + 1)] + d.VR[IX(1, J, K + 1)];
  float AVGPB1 = pe(g, 1, J, K) + pe(g, 1, J - 1, K) + pe(g, 1, J - 1, K + 1) + pe(g, 1, J, K + 1);
  float TF1 = .25f * (d.FLOWT[IX(1, J, K)] * AVGVR1 + AVGPB1 * d.ABR[IX(1, J, K)]);
  float AVGVRM = d.VR[IX(IM, J, K)] + d.VR[IX(IM, J - 1, K)] + d.VR[IX(IM, J - 1, K + 1)] + d.VR[IX(IM, J, K + 1)];
  float AVGPBM = pe(g, IM, J, K) + pe(g, IM, J - 1, K) + pe(g, IM, J - 1, K + 1) + pe(g, IM, J, K + 1);
  float TFM = .25f * (d.FLOWT[IX(IM, J, K)] * AVGVRM + AVGPBM * d.ABR[IX(IM, J, K)]);
  float AVGP1 = 0.25f * (pe(g, 1, J, K) + pe(g, 1, J - 1, K) + pe(g, 1, J, K + 1) + pe(g, 1, J - 1, K + 1));
  float AVGPM = 0.25f * (pe(g, IM, J, K) + pe(g, IM, J - 1, K) + pe(g, IM, J, K + 1) + pe(g, IM, J - 1, K + 1));
  float T1 = TF1 - AVGP1 * d.ABR[IX(1, J, K)];
  float T2 = TFM - AVGPM * d.ABR[IX(IM, J, K)];
  return 0.5f * ((T1 + T2) + (AVGP1 + AVGPM) * d.ABR[n]);
}

__device__ __forceinline__ void flux_radial_at(const Geo &g, const Dev &d, int I, int J, int K) {
  const size_t n = IX(I, J, K);
  if (I <= g.IMM1 && K <= g.KMM1) {
    float AVGVR = d.VR[IX(I, J, K)] + d.VR[IX(I, J, K + 1)] + d.VR[IX(I + 1, J, K)] + d.VR[IX(I + 1, J, K + 1)];
    float AVGP = d.PEFF[IX(I, J, K)] + d.PEFF[IX(I, J, K + 1)] + d.PEFF[IX(I + 1, J, K + 1)] + d.PEFF[IX(I + 1, J, K)];
    d.XFLUX[n] = 0.25f * (d.FLOWX[n] * AVGVR + d.AQR[n] * AVGP);
  }
  if (J >= 2 && K <= g.KMM1) {
    float AVGVR = d.VR[IX(I, J, K)] + d.VR[IX(I, J - 1, K)] + d.VR[IX(I, J - 1, K + 1)] + d.VR[IX(I, J, K + 1)];
    float AVGP = d.PEFF[IX(I, J, K)] + d.PEFF[IX(I, J - 1, K)] + d.PEFF[IX(I, J - 1, K + 1)] + d.PEFF[IX(I, J, K + 1)];
    float tf = .25f * (d.FLOWT[n] * AVGVR + AVGP * d.ABR[n]);
    if (d.IND[J - 1] != 1 && (I == 1 || I == g.IM)) tf = tflux_radial_per(g, d, n, J, K, PeffArray{d.PEFF});
    d.TFLUX[n] = tf;
  }
  if (J >= 2 && I <= g.IMM1) {
    float AVGVR = d.VR[IX(I, J, K)] + d.VR[IX(I, J - 1, K)] + d.VR[IX(I + 1, J - 1, K)] + d.VR[IX(I + 1, J, K)];
    float AVGP = d.PEFF[IX(I, J, K)] + d.PEFF[IX(I, J - 1, K)] + d.PEFF[IX(I + 1, J - 1, K)] + d.PEFF[IX(I + 1, J, K)];
    d.RFLUX[n] = 0.25f * (d.FLOWR[n] * AVGVR + d.ASR[n] * AVGP);
    if (K <= g.KMM1) {
      float AVGRVT = d.ROVT[IX(I, J, K)] + d.ROVT[IX(I, J, K + 1)] + d.ROVT[IX(I, J - 1, K)] + d.ROVT[IX(I, J - 1, K + 1)] +
                     d.ROVT[IX(I + 1, J, K)] + d.ROVT[IX(I + 1, J, K + 1)] + d.ROVT[IX(I + 1, J - 1, K)] +
                     d.ROVT[IX(I + 1, J - 1, K + 1)];
      float AVGVT = 0.125f * (d.VT[IX(I, J, K)] + d.VT[IX(I, J - 1, K)] + d.VT[IX(I, J - 1, K + 1)] + d.VT[IX(I, J, K + 1)] +
                              d.VT[IX(I + 1, J, K)] + d.VT[IX(I + 1, J - 1, K)] + d.VT[IX(I + 1, J - 1, K + 1)] +
                              d.VT[IX(I + 1, J, K + 1)]);
      float AVGP8 = d.PEFF[IX(I, J, K)] + d.PEFF[IX(I, J - 1, K)] + d.PEFF[IX(I, J - 1, K + 1)] + d.PEFF[IX(I, J, K + 1)] +
                    d.PEFF[IX(I + 1, J, K)] + d.PEFF[IX(I + 1, J - 1, K)] + d.PEFF[IX(I + 1, J - 1, K + 1)] +
                    d.PEFF[IX(I + 1, J, K + 1)];
      d.SOURCE[n] = d.RFORCE[n] - (AVGRVT * AVGVT + AVGP8) * d.VOLOR[n];
    }
  }
}
__global__ void k_flux_radial(Geo g, Dev d) {
  if (gated(d)) return;
  NODE_IJK
  flux_radial_at(g, d, I, J, K);
}

// The mass fluxes with the energy equation's fluxes (into e.q[0]) and, with
// nq = 2, the turbulence transport's (into e.q[1]) at the same node
// (continuity reads the mass-flux arrays themselves as its fluxes): the
// latter two take the node's mass fluxes from the thread instead of reading
// them back.  FLOWT(1,J,K), which the transport's pitchwise faces at I = IM
// read, is the node's own value on a periodic line and is formed again on a
// blade line (mf:3960-4190, 4357-4456, 4275-4335).
__device__ __forceinline__ FaceFlux sa_faces(const Geo &g, const Dev &d, int I, int J, int K, float fx, float ft,
                                             float fr, float ft1);
__global__ void k_massflux_eq(Geo g, Dev d, EqSet e, int nq) {
  NODE_IJK
  const int IM = g.IM, IMM1 = g.IMM1, KMM1 = g.KMM1;
  // the corners (I..I+1, J-1..J, K..K+1) with indices clamped into the
  // grid, every value loaded before anything waits: a clamped corner is
  // loaded but enters no face that is stored.  The expressions and their
  // operand order are those of massflux_vals, energy_faces and sa_faces;
  // only the periodic-line and blade-line ends (I = 1, IM) load more.
  const int Ip = min(I + 1, IM), Jm = max(J - 1, 1), Kp = min(K + 1, g.KM), Km = max(K - 1, 1);
  const unsigned n = IX(I, J, K);
  const unsigned c001 = IX(I, J, Kp), c100 = IX(Ip, J, K), c101 = IX(Ip, J, Kp);
  const unsigned cm00 = IX(I, Jm, K), cm01 = IX(I, Jm, Kp), cm10 = IX(Ip, Jm, K);
  const float x000 = d.ROVX[n], x001 = d.ROVX[c001], x100 = d.ROVX[c100], x101 = d.ROVX[c101], xm00 = d.ROVX[cm00],
              xm01 = d.ROVX[cm01], xm10 = d.ROVX[cm10];
  const float r000 = d.ROVR[n], r001 = d.ROVR[c001], r100 = d.ROVR[c100], r101 = d.ROVR[c101], rm00 = d.ROVR[cm00],
              rm01 = d.ROVR[cm01], rm10 = d.ROVR[cm10];
  const float w000 = d.ROVT[n], w001 = d.ROVT[c001], wm00 = d.ROVT[cm00], wm01 = d.ROVT[cm01];
  const float o000 = d.RO[n], o001 = d.RO[c001], om00 = d.RO[cm00], om01 = d.RO[cm01];
  const float u00 = d.UBLADE[JKX(J, K)], u01 = d.UBLADE[JKX(J, Kp)], um0 = d.UBLADE[JKX(Jm, K)],
              um1 = d.UBLADE[JKX(Jm, Kp)];
  const float h000 = d.HO[n], h001 = d.HO[c001], h100 = d.HO[c100], h101 = d.HO[c101], hm00 = d.HO[cm00],
              hm01 = d.HO[cm01], hm10 = d.HO[cm10];
  const float p000 = d.PEFF[n], p001 = d.PEFF[c001], pm00 = d.PEFF[cm00], pm01 = d.PEFF[cm01];
  const float aqx = d.AQX[n], aqr = d.AQR[n], abx = d.ABX[n], abr = d.ABR[n], asx = d.ASX[n], asr = d.ASR[n];
  const float abt = d.ABT[JKX(J, K)], wrabt = d.WRABT[JKX(J, K)], qs = d.QSOURCE[n];
  const bool sa = nq > 1;
  float tk0 = 0.f, tkjp = 0.f, tkkm = 0.f, tkim = 0.f, tsrc = 0.f;
  if (sa) {
    const float *TK = d.TRANS_KVIS;
    tk0 = TK[n];
    tkjp = TK[IX(I, min(J + 1, g.JM), K)];
    tkkm = TK[IX(I, J, Km)];
    tkim = TK[IX(max(I - 1, 1), J, K)];
    tsrc = d.T_SOURCE[n];
  }
  const int ind = d.IND[J - 1];
  const bool ends = I == 1 || I == IM;
  // the mass fluxes (massflux_vals)
  MassFlux m;
  if (I <= IMM1 && K <= KMM1) {
    float AVGRVX = x000 + x001 + x100 + x101;
    float AVGRVR = r000 + r001 + r100 + r101;
    m.fx = 0.25f * (AVGRVX * aqx + AVGRVR * aqr);
    m.hx = true;
  }
  m.rowt = w000 - u00 * o000;
  m.wt = m.rowt / o000;
  if (J >= 2 && K <= KMM1) {
    if (ind != 1 && ends) {
      m.ft = 0.5f * (flowt_raw(g, d, 1, J, K) + flowt_raw(g, d, IM, J, K));
    } else {
      float AVGRVX = x000 + xm00 + xm01 + x001;
      float AVGRVT = (w000 - u00 * o000) + (wm00 - um0 * om00) + (wm01 - um1 * om01) + (w001 - u01 * o001);
      float AVGRVR = r000 + rm00 + rm01 + r001;
      m.ft = (AVGRVX * abx + AVGRVT * abt + AVGRVR * abr) * 0.25f;
      if (ind == 1 && ends) {
        int NR = d.NROW[J - 1];
        if (K >= d.KS1[NR - 1] && K <= d.KS2[NR - 1]) m.ft = 0.0f;
      }
    }
    m.ht = true;
  }
  if (J >= 2 && I <= IMM1) {
    if (K >= 2 && K <= KMM1) {
      float AVGRVX = x000 + xm00 + xm10 + x100;
      float AVGRVR = r000 + rm00 + rm10 + r100;
      m.fr = 0.25f * (AVGRVX * asx + AVGRVR * asr);
    }
    m.hr = true;
  }
  // the energy equation's faces (energy_faces)
  FaceFlux fe;
  if (I <= IMM1 && K <= KMM1) {
    float AVGHO = h000 + h001 + h100 + h101;
    fe.x = m.fx * AVGHO * 0.25f;
    fe.hx = true;
    if (J >= 2) {
      fe.s = qs;
      fe.hs = true;
    }
  }
  if (J >= 2 && I <= IMM1) {
    float AVGHO = 0.25f * (h000 + hm00 + hm10 + h100);
    fe.r = m.fr * AVGHO;
    fe.hr = true;
  }
  if (J >= 2 && K <= KMM1) {
    if (ind != 1 && ends) {
      fe.t = 0.5f * (tflux_energy(g, d, 1, J, K, m.ft) + tflux_energy(g, d, IM, J, K, m.ft));
    } else {
      float AVGHO = h000 + hm00 + hm01 + h001;
      float AVGPB = p000 + pm00 + pm01 + p001;
      fe.t = (m.ft * AVGHO + wrabt * AVGPB) * 0.25f;
    }
    fe.ht = true;
  }
  // the transport's faces (sa_faces); FLOWT(1,J,K), which the pitchwise
  // face at I = IM reads, is the node's own value on a periodic line and is
  // formed again on a blade line
  FaceFlux fs;
  if (sa) {
    if (K == g.KM) {
      fs.hr = J >= 2 && I <= IMM1;
    } else {
      if (I <= IMM1) {
        const bool jp = J != g.JM && J != 1 && d.INDMIX[J - 1] != 1;  // JP1 = J + 1, else J
        fs.x = m.fx * (tk0 + (jp ? tkjp : tk0));
        fs.hx = true;
      }
      if (J >= 2) {
        if (I <= IMM1) {
          fs.s = -2.0f * tsrc;
          fs.r = (K == 1) ? 0.0f : m.fr * (tkkm + tk0);
          fs.hs = fs.hr = true;
        }
        if (I >= 2 && I <= IMM1) {
          fs.t = m.ft * (tkim + tk0);
          fs.ht = true;
        }
        if (ends) {
          float ft1 = m.ft;
          if (I == IM && K <= KMM1 && ind == 1) {
            ft1 = flowt_raw(g, d, 1, J, K);
            const int NR = d.NROW[J - 1];
            if (K >= d.KS1[NR - 1] && K <= d.KS2[NR - 1]) ft1 = 0.0f;
          }
          fs.t = ft1 * (d.TRANS_KVIS[IX(IMM1, J, K)] + d.TRANS_KVIS[IX(1, J, K)]);
          fs.ht = true;
        }
      }
    }
  }
  if (gated(d)) return;
  massflux_store(g, d, m, n, false);
  store_faces(fe, e.q[0].XFLUX, e.q[0].RFLUX, e.q[0].TFLUX, e.q[0].SOURCE, n);
  if (sa) store_faces(fs, e.q[1].XFLUX, e.q[1].RFLUX, e.q[1].TFLUX, e.q[1].SOURCE, n);
}

// The three momentum equations' face fluxes and sources at one node in one
// thread (axial mf:4818-4896, angular mf:4913-4956, radial mf:5003-5095),
// into the flux arrays of e.q[0], e.q[1] and e.q[2]: the node values they
// share (PEFF at the eight corners, the mass fluxes) are loaded once and
// every value before any store.  The expressions and their operand order
// are those of k_flux_axial, k_flux_angular and k_flux_radial.  PEFF
// (mf:4751-4757) is formed at the corners from P as k_peff forms it (P at
// the corner and at the next J, the last J taking itself) and stored at the
// thread's own node, so the step needs no separate PEFF launch.
__device__ __forceinline__ float peff_of(const Dev &d, const Params &p, const StepScalars &s, unsigned a,
                                         unsigned b) {
  return d.P[a] - s.PREFF + p.F_PDOWN * (d.P[b] - d.P[a]);
}
__global__ void k_flux_mom(Geo g, Dev d, EqSet e, Params pr, StepScalars ss) {
  NODE_IJK
  const int IM = g.IM;
  const bool fx = I <= g.IMM1 && K <= g.KMM1, fr = J >= 2 && I <= g.IMM1, ft = J >= 2 && K <= g.KMM1;
  const bool per = ft && d.IND[J - 1] != 1 && (I == 1 || I == IM);
  // the corners (I..I+1, J-1..J, K..K+1) with indices clamped into the
  // grid: a clamped corner is loaded but enters no face that is stored
  const int Ip = min(I + 1, IM), Jm = max(J - 1, 1), Kp = min(K + 1, g.KM);
  const unsigned n = IX(I, J, K);
  const unsigned c001 = IX(I, J, Kp), c100 = IX(Ip, J, K), c101 = IX(Ip, J, Kp);
  const unsigned cm00 = IX(I, Jm, K), cm01 = IX(I, Jm, Kp), cm10 = IX(Ip, Jm, K), cm11 = IX(Ip, Jm, Kp);
#define C7(A, v)                                                                                            \
  const float v##000 = d.A[n], v##001 = d.A[c001], v##100 = d.A[c100], v##101 = d.A[c101], v##m00 = d.A[cm00], \
              v##m01 = d.A[cm01], v##m10 = d.A[cm10];
#define C8(A, v) C7(A, v) const float v##m11 = d.A[cm11];
  // the corners' next-J nodes: planes J+1 (J at the last plane) and J
  const int Jn = (J == g.JM) ? J : J + 1;
  const unsigned n001 = IX(I, Jn, Kp), n100 = IX(Ip, Jn, K), n101 = IX(Ip, Jn, Kp), n000 = IX(I, Jn, K);
  const float p000 = peff_of(d, pr, ss, n, n000), p001 = peff_of(d, pr, ss, c001, n001),
              p100 = peff_of(d, pr, ss, c100, n100), p101 = peff_of(d, pr, ss, c101, n101),
              pm00 = peff_of(d, pr, ss, cm00, n), pm01 = peff_of(d, pr, ss, cm01, c001),
              pm10 = peff_of(d, pr, ss, cm10, c100), pm11 = peff_of(d, pr, ss, cm11, c101);
  C7(VX, x)
  C8(VT, t)
  C7(VR, r)
  C8(ROVT, w)
#undef C8
#undef C7
  const float flx = d.FLOWX[n], flt = d.FLOWT[n], flr = d.FLOWR[n];
  const float R0 = d.R[JKX(J, K)], R1 = d.R[JKX(J, Kp)], Rm = d.R[JKX(Jm, K)];
  float axX = 0.f, axS = 0.f, anX = 0.f, anS = 0.f, raX = 0.f;
  if (fx) {
    const float AVGP = p000 + p001 + p101 + p100;
    const float AVGVX = x000 + x001 + x101 + x100;
    axS = d.XFORCE[n];
    axX = 0.25f * (AVGP * d.AQX[n] + flx * AVGVX);
    const float AVGR = 0.125f * (R0 + R1);
    const float AVGVT = t000 + t001 + t100 + t101;
    anS = d.TFORCE[n];
    anX = flx * AVGR * AVGVT;
    const float AVGVR = r000 + r001 + r100 + r101;
    raX = 0.25f * (flx * AVGVR + d.AQR[n] * AVGP);
  }
  float axR = 0.f, anR = 0.f, raR = 0.f, raS = 0.f;
  if (fr) {
    const float AVGP = p000 + pm00 + pm10 + p100;
    const float AVGVX = x000 + xm00 + xm10 + x100;
    axR = 0.25f * (AVGP * d.ASX[n] + flr * AVGVX);
    const float AVGR = (R0 + Rm) * 0.125f;
    const float AVGVT = t000 + tm00 + tm10 + t100;
    anR = flr * AVGVT * AVGR;
    const float AVGVR = r000 + rm00 + rm10 + r100;
    raR = 0.25f * (flr * AVGVR + d.ASR[n] * AVGP);
    if (K <= g.KMM1) {
      const float AVGRVT = w000 + w001 + wm00 + wm01 + w100 + w101 + wm10 + wm11;
      const float AVGVT8 = 0.125f * (t000 + tm00 + tm01 + t001 + t100 + tm10 + tm11 + t101);
      const float AVGP8 = p000 + pm00 + pm01 + p001 + p100 + pm10 + pm11 + p101;
      raS = d.RFORCE[n] - (AVGRVT * AVGVT8 + AVGP8) * d.VOLOR[n];
    }
  }
  float axT = 0.f, anT = 0.f, raT = 0.f;
  if (ft) {
    const float AVGP = p000 + pm00 + pm01 + p001;
    if (per) {
      const PeffFromP pe{d.P, ss.PREFF, pr.F_PDOWN};
      axT = tflux_axial_per(g, d, n, J, K, pe);
      anT = 0.5f * (tflux_angular(g, d, 1, J, K, pe) + tflux_angular(g, d, IM, J, K, pe));
      raT = tflux_radial_per(g, d, n, J, K, pe);
    } else {
      const float AVGVX = x000 + xm00 + xm01 + x001;
      axT = 0.25f * (flt * AVGVX + d.ABX[n] * AVGP);
      const float AVGR = 0.25f * d.RAVG_CELL[JKX(J, K)];
      const float AVGVT = t000 + tm00 + tm01 + t001;
      anT = (flt * AVGVT + d.ABT[JKX(J, K)] * AVGP) * AVGR;
      const float AVGVR = r000 + rm00 + rm01 + r001;
      raT = .25f * (flt * AVGVR + AVGP * d.ABR[n]);
    }
  }
  const EqLane &qa = e.q[0], &qt = e.q[1], &qr = e.q[2];
  if (gated(d)) return;
  d.PEFF[n] = p000;
  if (fx) {
    qa.SOURCE[n] = axS;
    qa.XFLUX[n] = axX;
    qt.SOURCE[n] = anS;
    qt.XFLUX[n] = anX;
    qr.XFLUX[n] = raX;
  }
  if (fr) {
    qa.RFLUX[n] = axR;
    qt.RFLUX[n] = anR;
    qr.RFLUX[n] = raR;
    if (K <= g.KMM1) qr.SOURCE[n] = raS;
  }
  if (ft) {
    qa.TFLUX[n] = axT;
    qt.TFLUX[n] = anT;
    qr.TFLUX[n] = raT;
  }
}

// ---------------------------------------------------------------------------
// inlet and exit boundary conditions, mf:4476-4734 (IFGAS=0, ITIMST<5,
// no throttle or plate loss).  k_bc_lines forms the pitchwise sums per K,
// k_bc_pd the exit-pressure recursion PD(K), which is serial in K, and
// k_bc_apply updates the inlet and exit nodes in parallel.  d.BCW carries
// PAVG2, PAVG3 and PAVGOUT (KM each) between them.

// inlet pressure and static temperature of node (I,1,K) from the state
// before the update, mf:4566-4606

__device__ __forceinline__ void inlet_pt(const Geo &g, const Dev &d, const Params &p, const StepScalars &s,
                                         const float *PAVG2, const float *PAVG3, int I, int K, float &P1,
                                         float &TSTATIC) {
  const unsigned n1 = IX(I, 1, K);
  const float PO1 = d.PO1[K - 1], TO1 = d.TO1[K - 1];
  float RO_STAG = PO1 / p.RGAS / TO1;
  float RO_IN = d.RO[n1];
  if (RO_IN > 0.9999f * RO_STAG) RO_IN = 0.9999f * RO_STAG;
  float PNEW = PO1 * pow_r(RO_IN / RO_STAG, p.GA);
  P1 = d.P[n1];
  if (p.IN_PRESS == 0) P1 = s.RFIN1 * P1 + p.RFIN * PNEW;
  if (abs(p.IN_PRESS) == 1) P1 = s.RFIN1 * P1 + s.RFIN2 * d.P[IX(I, 2, K)] - p.RFIN * d.P[IX(I, 3, K)];
  if (abs(p.IN_PRESS) == 3) P1 = s.RFIN1 * P1 + s.RFIN2 * PAVG2[K - 1] - p.RFIN * PAVG3[K - 1];
  if (abs(p.IN_PRESS) == 4)
    P1 = s.RFIN1 * P1 + s.RFIN2 * d.P[IX(p.IMID, 2, p.KMID)] - p.RFIN * d.P[IX(p.IMID, 3, p.KMID)];
  const float plim = d.SC[SC_PLIM + K - 1];
  if (P1 > plim) P1 = 0.9f * P1 + 0.1f * plim;
  if (P1 > 0.99999f * PO1) P1 = PO1 * 0.99999f;
  TSTATIC = TO1 * pow_r(P1 * d.RPO1[K - 1], p.FGA);
}

// Pitchwise sums per K in the reference's order over I, mf:4512-4521: the
// static pressure averages on J = 2, 3 (inlet) and JM-1 (exit) and the
// swirl term of the exit's radial equilibrium.  One wavefront per K line,
// into d.BCW (PAVG2, PAVG3, PAVGOUT at 0, KM, 2 KM; the swirl sums at
// 3 KM + 4).
__global__ void __launch_bounds__(64) k_bc_lines(Geo g, Dev d, Params p) {
  if (gated(d)) return;
  const int K = blockIdx.x + 1, lane = threadIdx.x;
  const int JM = g.JM, JMM1 = g.JMM1, IMM1 = g.IMM1;
  float v2[4], v3[4], vo[4], vs[4];
  const float rjm = d.R[JKX(JM, K)];
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int I = 64 * c + lane + 1;
    if (I <= IMM1) {
      const float fp = d.FP[I - 1];
      v2[c] = 0.5f * (d.P[IX(I, 2, K)] + d.P[IX(I + 1, 2, K)]) * fp;
      v3[c] = 0.5f * (d.P[IX(I, 3, K)] + d.P[IX(I + 1, 3, K)]) * fp;
      vo[c] = 0.5f * (d.P[IX(I, JMM1, K)] + d.P[IX(I + 1, JMM1, K)]) * fp;
      vs[c] = 0.5f * fp *
              (d.ROVT[IX(I, JM, K)] * d.VT[IX(I, JM, K)] + d.ROVT[IX(I + 1, JM, K)] * d.VT[IX(I + 1, JM, K)]) / rjm;
    } else {
      v2[c] = v3[c] = vo[c] = vs[c] = 0.0f;
    }
  }
  const float a2 = line_sum(v2, IMM1, p.EXACT), a3 = line_sum(v3, IMM1, p.EXACT);
  const float ao = line_sum(vo, IMM1, p.EXACT), sm = line_sum(vs, IMM1, p.EXACT);
  if (lane == 0) {
    d.BCW[K - 1] = a2;
    d.BCW[g.KM + K - 1] = a3;
    d.BCW[2 * g.KM + K - 1] = ao;
    d.BCW[3 * g.KM + 4 + K - 1] = sm;
  }
}

// Exit pressure from radial equilibrium, serial in K (mf:4531-4542).  The
// workgroup stages the per-K inputs in LDS; thread 0 runs the recursion
// with the running value in a register, reading only LDS, while thread 64
// forms the SA inlet density of the mid-passage node.  The reference resets
// PD(1) to PDOWN_HUB on every pass of its K loop, so after the loop PD(1)
// is PDOWN_HUB and the K = 2 update reads it without the IPOUT = -1 shift;
// but the exit nodes of K = 1 are set inside the K = 1 pass, when PD(1)
// still holds PDOWN_HUB + PDIFF.  That value goes to d.BCW[3 KM + 2] for
// k_bc_apply.
constexpr int kBcMaxK = 512;
__global__ void __launch_bounds__(128) k_bc_pd(Geo g, Dev d, Params p, StepScalars s) {
  if (gated(d)) return;
  __shared__ float sumk[kBcMaxK], rjm[kBcMaxK], pdold[kBcMaxK], pdnew[kBcMaxK];
  const float *PAVG2 = d.BCW, *PAVG3 = d.BCW + g.KM, *SUMK = d.BCW + 3 * g.KM + 4;
  const int KM = g.KM, JM = g.JM;
  for (int K = threadIdx.x + 1; K <= KM; K += blockDim.x) {
    sumk[K - 1] = SUMK[K - 1];
    rjm[K - 1] = d.R[JKX(JM, K)];
    pdold[K - 1] = d.PD[K - 1];
  }
  __syncthreads();
  if (threadIdx.x == 0 && p.IPOUT <= 0) {
    float PDIFF = 0.0f;
    if (p.IPOUT == -1) PDIFF = p.PDOWN_TIP - pdold[KM - 1];
    float prev = p.PDOWN_HUB, SUMP = sumk[0];  // PD(K-1) and SUMK(K-1) as the K-th pass reads them
#pragma unroll 8
    for (int K = 2; K <= KM; K++) {
      const float DP = (sumk[K - 1] + SUMP) * 0.5f * (rjm[K - 1] - rjm[K - 2]);
      float v = 0.9f * pdold[K - 1] + 0.1f * (prev + DP);
      if (p.IPOUT == -1) v = v + PDIFF;
      pdnew[K - 1] = v;
      SUMP = sumk[K - 1];
      prev = v;
    }
    pdnew[0] = (KM >= 2) ? p.PDOWN_HUB : p.PDOWN_HUB + PDIFF;
    d.BCW[3 * KM + 2] = p.PDOWN_HUB + PDIFF;
  } else if (threadIdx.x == 0) {
    d.BCW[3 * KM + 2] = pdold[0];
  }
  if (threadIdx.x == 64 && p.ILOS >= 200) {
    // RO(IMID,1,KMID) before and after its inlet update (the SA inlet
    // condition reads whichever the reference's loop has reached)
    float P1, TSTATIC;
    inlet_pt(g, d, p, s, PAVG2, PAVG3, p.IMID, p.KMID, P1, TSTATIC);
    d.BCW[3 * g.KM] = d.RO[IX(p.IMID, 1, p.KMID)];
    d.BCW[3 * g.KM + 1] = P1 / (p.RGAS * TSTATIC);
  }
  __syncthreads();
  if (p.IPOUT <= 0)
    for (int K = threadIdx.x + 1; K <= KM; K += blockDim.x) d.PD[K - 1] = pdnew[K - 1];
}

__global__ void k_bc_apply(Geo g, Dev d, Params p, StepScalars s) {
  if (gated(d)) return;
  const float *PAVG2 = d.BCW, *PAVG3 = d.BCW + g.KM, *PAVGOUT = d.BCW + 2 * g.KM, *PDL = d.PD;
  const int JM = g.JM, JMM1 = g.JMM1, IM = g.IM;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= IM * g.KM) return;
  {
    const int I = t % IM + 1, K = t / IM + 1;
    const size_t n1 = IX(I, 1, K);
    const float TO1 = d.TO1[K - 1];
    float P1, TSTATIC;
    inlet_pt(g, d, p, s, PAVG2, PAVG3, I, K, P1, TSTATIC);
    d.P[n1] = P1;
    float VABSQ = 2 * p.CP * (TO1 - TSTATIC);
    d.T_STATIC[n1] = TSTATIC;
    float VABS = sqrtf(VABSQ);
    float VMER = 0.0f, vt1 = d.VT[n1];
    if (p.IN_VTAN == 0) {
      VMER = VABS * d.BSCOS[K - 1];
      vt1 = VABS * d.BSSIN[K - 1];
    }
    if (p.IN_VTAN == 1) {
      float VMERSQ = VABSQ - d.VTIN[K - 1] * d.VTIN[K - 1];
      if (VMERSQ < 0.01f) VMERSQ = 0.01f;
      VMER = sqrtf(VMERSQ);
      vt1 = d.VTIN[K - 1];
    }
    if (p.IN_VTAN == 2) {
      float WR = d.UBLADE[JKX(1, K)];
      float VRELSQ = VABSQ - WR * WR * d.BSCOS[K - 1] * d.BSCOS[K - 1];
      if (VRELSQ < 0.01f) VRELSQ = 0.01f;
      float VREL = sqrtf(VRELSQ) - WR * d.BSSIN[K - 1];
      VMER = VREL * d.BSCOS[K - 1];
      vt1 = VREL * d.BSSIN[K - 1] + WR;
    }
    float vx1 = d.VX[n1], vr1 = d.VR[n1];
    const float vr2 = d.VR[IX(I, 2, K)];
    if (p.IN_VR > 0) vx1 = VMER * d.BRCOS[K - 1];
    float VXSQ = VMER * VMER - vr2 * vr2;
    if (VXSQ < 0.0000001f * TO1 * p.CP) VXSQ = 0.0000001f * TO1 * p.CP;
    if (p.IN_VR <= 0) vx1 = sqrtf(VXSQ);
    if (p.IN_VR > 0) vr1 = VMER * d.BRSIN[K - 1];
    if (p.IN_VR <= 0) vr1 = vr2;
    float ro1 = P1 / (p.RGAS * TSTATIC);
    float rovt1 = ro1 * vt1;
    d.VX[n1] = vx1;
    d.VR[n1] = vr1;
    d.VT[n1] = vt1;
    d.RO[n1] = ro1;
    d.ROVT[n1] = rovt1;
    d.ROVR[n1] = ro1 * vr1;
    d.RORVT[n1] = rovt1 * d.R[JKX(1, K)];
    d.ROVX[n1] = ro1 * vx1;
    d.ROE[n1] = P1 / p.GA1 + ro1 * 0.5f * VABSQ;
    d.HO[n1] = p.CP * TO1;
    if (p.ILOS >= 200) {
      // inlet turbulent viscosity, mf:4679-4690; the reference divides by
      // RO(IMID,1,KMID) as its (K, I) loop has left it: already updated
      // for (K,I) at or after (KMID,IMID), k_bc_pd kept both values
      const float FST = d.FSTURB[0], VL = d.VISC_LAM[IX(I, 2, K)];
      const float TRANSVISIN = (FST < 1.0f) ? VL * 4.35f * pow_r(FST, 0.25f) : VL * (3.5f + FST * 0.85f);
      const bool after = K > p.KMID || (K == p.KMID && I >= p.IMID);
      const float roref = after ? d.BCW[3 * g.KM + 1] : d.BCW[3 * g.KM];
      d.TRANS_DYN_VIS[n1] = TRANSVISIN;
      d.TRANS_DYN_VIS[IX(I, 2, K)] = TRANSVISIN;
      const float tk = TRANSVISIN / roref;
      d.TRANS_KVIS[n1] = tk;
      d.TRANS_KVIS[IX(I, 2, K)] = tk;
    }
    // exit, mf:4690-4730
    const size_t nm = IX(I, JM, K);
    float vr = d.VR[nm], vt = d.VT[nm], vx = d.VX[nm];
    float EKE = .5f * (vr * vr + vt * vt + vx * vx);
    const float pdk = K == 1 ? d.BCW[3 * g.KM + 2] : PDL[K - 1];  // PD(1) as the K = 1 pass leaves it
    float PEX = pdk + p.FEXTRAP * (d.P[IX(I, JMM1, K)] - PAVGOUT[K - 1]);
    float HOEX = 2.0f * d.HO[IX(I, JMM1, K)] - d.HO[IX(I, JM - 2, K)];
    float TEXIT = (HOEX - EKE) / p.CP;
    float roex = PEX / p.RGAS / TEXIT;
    d.P[nm] = PEX;
    d.HO[nm] = HOEX;
    d.RO[nm] = roex;
    d.ROE[nm] = roex * (p.CV * TEXIT + EKE);
  }
}

// PEFF, mf:4751-4757
__global__ void k_peff(Geo g, Dev d, Params p, StepScalars s) {
  NODE_IJK
  int JP1 = J + 1;
  if (J == g.JM) JP1 = g.JM;
  const size_t n = IX(I, J, K);
  const float v = d.P[n] - s.PREFF + p.F_PDOWN * (d.P[IX(I, JP1, K)] - d.P[n]);
  if (gated(d)) return;
  d.PEFF[n] = v;
}

// ---------------------------------------------------------------------------
// TSTEP phases (mf:5470-5932)

// upstream-face flux extrapolation at mixing planes, mf:5516-5555
__global__ void k_fextrap(Geo g, Dev d, Params p) {
  if (gated(d)) return;
  const int t = blockIdx.x;  // one wavefront per (mixing plane, K) line
  const int NR = t / g.KMM1 + 1, K = t % g.KMM1 + 1;
  if (NR > g.NROWS - 1) return;
  const int lane = threadIdx.x;
  const int J = d.JMIX[NR - 1], JP1 = J + 1, JM1 = J - 1, JP2 = J + 2;
  const float FLOWDIRN = -d.FLOWX[IX(p.IMID, J, K)];
  float a[4], b[4];
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int I = 64 * c + lane + 1;
    a[c] = (I <= g.IMM1) ? d.XFLUX[IX(I, JM1, K)] : 0.0f;
    b[c] = (I <= g.IMM1) ? d.XFLUX[IX(I, JP2, K)] : 0.0f;
  }
  const float AVFLXJM1 = line_sum(a, g.IMM1, p.EXACT);
  const float AVFLXJP2 = line_sum(b, g.IMM1, p.EXACT);
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int I = 64 * c + lane + 1;
    if (I > g.IMM1) continue;
    if (FLOWDIRN > 0.0f) {
      float DFLUX = (a[c] - AVFLXJM1 * d.FP[I - 1]) * p.FEXTRAP;
      d.SOURCE[IX(I, J, K)] = d.SOURCE[IX(I, J, K)] - DFLUX;
    } else {
      float DFLUX = (b[c] - AVFLXJP2 * d.FP[I - 1]) * p.FEXTRAP;
      d.SOURCE[IX(I, JP2, K)] = d.SOURCE[IX(I, JP2, K)] + DFLUX;
    }
    d.SOURCE[IX(I, JP1, K)] = 0.0f;
  }
}

// flux sum, scree/SSS history (mf:5560-5570), with the tip-gap TFLUX
// balance of mf:5503-5510 applied to the two periodic faces it touches.
__global__ void k_residual(Geo g, Dev d, float *D, float *DIFF, Params p) {
  if (gated(d)) return;
  NODE_IJK
  if (I > g.IMM1 || K > g.KMM1 || J < 2) return;
  const size_t n = IX(I, J, K);
  float RATPITCH = (float)d.NBLADE[J - 2] / (float)d.NBLADE[J - 1];
  float TF = d.TFLUX[n], TFP = d.TFLUX[IX(I + 1, J, K)];
  const int NR = d.NROW[J - 1];
  const int kts = d.KTIPS[NR - 1];
  if (kts > 0 && K >= kts && K <= d.KTIPE[NR - 1] - 1) {
    float tavg = 0.5f * (d.TFLUX[IX(1, J, K)] + d.TFLUX[IX(g.IM, J, K)]);
    if (I == 1) TF = tavg;
    if (I + 1 == g.IM) TFP = tavg;
  }
  float DELTA = d.XFLUX[n] - d.XFLUX[IX(I, J - 1, K)] * RATPITCH + d.RFLUX[n] - d.RFLUX[IX(I, J, K + 1)] + TF - TFP -
                d.SOURCE[n];
  float df = DIFF[n];
  d.STORE[n] = p.F1 * DELTA + p.F2 * df;
  DIFF[n] = DELTA + p.F3 * df;
}

// pitchwise averaging of STORE at the mixing planes, mf:5578-5607 (pass 1,
// FP-weighted) and mf:5786-5818 (pass 2, plain mean)
__global__ void k_mpavg(Geo g, Dev d, Params p, int pass) {
  if (gated(d)) return;
  const int t = blockIdx.x;  // one wavefront per (mixing plane, K) line
  const int NR = t / g.KMM1 + 1, K = t % g.KMM1 + 1;
  if (NR > g.NROWS - 1) return;
  const int lane = threadIdx.x;
  const int J = d.JMIX[NR - 1];
  const float FLOWDIRN = -d.FLOWX[IX(p.IMID, J, K)];
  for (int side = 0; side < 2; side++) {
    const bool on = side == 0 ? ((FLOWDIRN > 0.0f) || (p.FEXTRAP < 0.001f))
                              : ((FLOWDIRN < 0.0f) || (p.FEXTRAP < 0.001f));
    if (!on) continue;
    const int JMX = side == 0 ? J + 2 : J;
    float v[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int I = 64 * c + lane + 1;
      v[c] = (I <= g.IMM1) ? d.STORE[IX(I, JMX, K)] : 0.0f;
    }
    float SUM_STORE = line_sum(v, g.IMM1, p.EXACT);
    if (pass != 1) SUM_STORE = SUM_STORE / (float)g.IMM1;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int I = 64 * c + lane + 1;
      if (I <= g.IMM1) d.STORE[IX(I, JMX, K)] = (pass == 1) ? SUM_STORE * d.FP[I - 1] : SUM_STORE;
    }
  }
}

// Multigrid block sums, mf:5621-5635.  The J members of each block come
// from a host-built CSR list (jptr/jlist) because blocks stop at mixing
// planes and the mixing-plane cells form a separate dummy block.
// Level 1 (3x3x3 cells): one thread per block, reference order K, J, I.
__global__ void k_mg_sum_seq(Geo g, Dev d, const int *jptr, const int *jlist, int ni, int nj, int nk, int ir, int kr,
                             float *out) {
  if (gated(d)) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ni * nj * nk) return;
  const int ib = t % ni + 1, jb = (t / ni) % nj + 1, kb = t / (ni * nj) + 1;
  const int i0 = (ib - 1) * ir + 1, i1 = min(ib * ir, g.IMM1);
  const int k0 = (kb - 1) * kr + 1, k1 = min(kb * kr, g.KMM1);
  float s = 0.0f;
  for (int K = k0; K <= k1; K++)
    for (int q = jptr[jb - 1]; q < jptr[jb]; q++) {
      const int J = jlist[q];
      for (int I = i0; I <= i1; I++) s = s + d.STORE[IX(I, J, K)];
    }
  out[t] = s;
}

// Level 2 (9x9x9) in FP64: one wavefront per block, lanes stride over the
// block's cells, fixed-order tree at the end (deterministic).
__global__ void k_mg_sum_wave(Geo g, Dev d, const int *jptr, const int *jlist, int ni, int nj, int nk, int ir, int kr,
                              float *out) {
  if (gated(d)) return;
  const int t = blockIdx.x;
  if (t >= ni * nj * nk) return;
  const int ib = t % ni + 1, jb = (t / ni) % nj + 1, kb = t / (ni * nj) + 1;
  const int i0 = (ib - 1) * ir + 1, i1 = min(ib * ir, g.IMM1);
  const int k0 = (kb - 1) * kr + 1, k1 = min(kb * kr, g.KMM1);
  const int ni_ = i1 - i0 + 1, nk_ = k1 - k0 + 1, q0 = jptr[jb - 1], nj_ = jptr[jb] - q0;
  const int tot = ni_ * nj_ * nk_;
  double s = 0.0;
  for (int c = threadIdx.x; c < tot; c += 64) {
    const int I = i0 + c % ni_, J = jlist[q0 + (c / ni_) % nj_], K = k0 + c / (ni_ * nj_);
    s += (double)d.STORE[IX(I, J, K)];
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (threadIdx.x == 0) out[t] = (float)s;
}

// Superblock and per-row sums over whole J ranges.  Stage 1: one
// workgroup per J plane sums its cells (optionally |STORE|) in FP64 with a
// fixed-order tree.
__global__ void k_jplane_sums(Geo g, Dev d, double *jpart, int absval) {
  if (gated(d)) return;
  __shared__ double sh[256];
  const int J = blockIdx.x + 1;
  double s = 0.0;
  if (J >= 2) {
    const int n = g.IMM1 * g.KMM1;
    for (int c = threadIdx.x; c < n; c += 256) {
      const int I = c % g.IMM1 + 1, K = c / g.IMM1 + 1;
      const float v = d.STORE[IX(I, J, K)];
      s += (double)(absval ? fabsf(v) : v);
    }
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) jpart[J - 1] = sh[0];
}

// stage 2: one wavefront per J range sums the per-J partials
__global__ void k_range_sums(Geo g, const double *jpart, const int *jptr, const int *jlist, int nranges, float *out,
                             int divide_by_cells, const int *gate) {
  if (gate && *gate) return;
  const int r = blockIdx.x;
  if (r >= nranges) return;
  const int q0 = jptr[r], nj = jptr[r + 1] - q0;
  double s = 0.0;
  for (int c = threadIdx.x; c < nj; c += 64) s += jpart[jlist[q0 + c] - 1];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (threadIdx.x == 0) {
    float f = (float)s;
    if (divide_by_cells) f = f / (float)(g.IMM1 * g.KMM1 * nj);
    out[r] = f;
  }
}

// sequential-order versions for EXACT mode (reference order K, J, I)
__global__ void k_range_sums_seq(Geo g, Dev d, const int *jptr, const int *jlist, int nranges, float *out,
                                 int absval, int divide_by_cells) {
  if (gated(d)) return;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nranges) return;
  float s = 0.0f;
  for (int K = 1; K <= g.KMM1; K++)
    for (int q = jptr[r]; q < jptr[r + 1]; q++) {
      const int J = jlist[q];
      for (int I = 1; I <= g.IMM1; I++) {
        float v = d.STORE[IX(I, J, K)];
        s = s + (absval ? fabsf(v) : v);
      }
    }
  if (divide_by_cells) s = s / (float)(g.IMM1 * g.KMM1 * (jptr[r + 1] - jptr[r]));
  out[r] = s;
}

// multigrid update, mf:5645-5660
__global__ void k_mg_apply(Geo g, Dev d) {
  if (gated(d)) return;
  NODE_IJK
  if (I > g.IMM1 || K > g.KMM1 || J < 2) return;
  const size_t n = IX(I, J, K);
  const int I1 = d.IB1[I - 1], I2 = d.IB2[I - 1], K1 = d.KB1[K - 1], K2 = d.KB2[K - 1];
  const int J1 = d.JB1[J - 1], J2 = d.JB2[J - 1], JSB = d.JSBLK[J - 1];
  const size_t b1 = (size_t)(I1 - 1) + (size_t)g.NIB1 * ((size_t)(J1 - 1) + (size_t)(g.NJB1 + 1) * (K1 - 1));
  const size_t b2 = (size_t)(I2 - 1) + (size_t)g.NIB2 * ((size_t)(J2 - 1) + (size_t)(g.NJB2 + 1) * (K2 - 1));
  float DELTA = d.STORE[n] * d.STEP[n] +
                (d.B1CHG[b1] * d.STEP1[b1] + d.B2CHG[b2] * d.STEP2[b2] + d.SBCHG[JSB - 1] * d.STEPSBK[JSB - 1]) * d.RSTEP[n];
  d.STORE[n] = DELTA;
}

// SMOOTH_RESID, mf:14833-14917: one pass in J, then I, then K, within each
// row (rows never share cells).  dir 0/1/2 = J/I/K; src -> dst.
// SMFAC is the smoothing factor, times the step's SFT when sft is set
__global__ void k_smooth_resid(Geo g, Dev d, const float *src, float *dst, float SMFAC, int sft, int dir) {
  if (gated(d)) return;
  if (sft) SMFAC = SMFAC * d.SC[SC_SFT];
  NODE_IJK
  if (I > g.IMM1 || K > g.KMM1) return;
  const int NR = d.NROW[J - 1];
  const int J1 = d.JSTART[NR - 1], J1P1 = J1 + 1, J1P2 = J1 + 2, J2 = d.JMIX[NR - 1], J2M1 = J2 - 1;
  const size_t n = IX(I, J, K);
  if (J < J1P1) { dst[n] = src[n]; return; }
  const float SFAC1 = 1.0f - SMFAC, SFACH = 0.5f * SMFAC;
  float v;
  if (dir == 0) {
    if (J == J1P1) v = src[n] + SFACH * (src[IX(I, J1P2, K)] - src[n]);
    else if (J == J2) v = src[n] + SFACH * (src[IX(I, J2M1, K)] - src[n]);
    else v = SFAC1 * src[n] + SFACH * (src[IX(I, J + 1, K)] + src[IX(I, J - 1, K)]);
  } else if (dir == 1) {
    const int IMM1 = g.IMM1, IMM2 = g.IM - 2;
    if (I == 1) v = src[n] + SFACH * (src[IX(2, J, K)] - src[n]);
    else if (I == IMM1) v = src[n] + SFACH * (src[IX(IMM2, J, K)] - src[n]);
    else v = SFAC1 * src[n] + SFACH * (src[IX(I + 1, J, K)] + src[IX(I - 1, J, K)]);
  } else {
    const int KMM1 = g.KMM1, KMM2 = g.KM - 2;
    if (K == 1) v = src[n] + SFACH * (src[IX(I, J, 2)] - src[n]);
    else if (K == KMM1) v = src[n] + SFACH * (src[IX(I, J, KMM2)] - src[n]);
    else v = SFAC1 * src[n] + SFACH * (src[IX(I, J, K + 1)] + src[IX(I, J, K - 1)]);
  }
  dst[n] = v;
}

// row-averaged change smoothing between neighbouring rows (mf:5699-5710)
// evaluated per thread for its own row, then the negative feedback
// (mf:5716-5724)
__device__ __forceinline__ float avg_blk(const float *AC, int NR, int NROWS) {
  if (NROWS == 1) return AC[0];
  if (NROWS == 2) return 0.5f * (AC[0] + AC[1]);
  if (NR == 1) return 0.5f * (AC[0] + AC[1]);
  if (NR == NROWS) return 0.5f * (AC[NROWS - 2] + AC[NROWS - 1]);
  return 0.25f * (AC[NR - 2] + AC[NR]) + 0.5f * AC[NR - 1];
}

__global__ void k_damp_apply(Geo g, Dev d) {
  if (gated(d)) return;
  const float DAMP = d.SC[SC_DAMP];
  NODE_IJK
  if (I > g.IMM1 || K > g.KMM1 || J < 2) return;
  const size_t n = IX(I, J, K);
  const int NR = d.NROW[J - 1];
  float DELTA = d.STORE[n];
  float ABSCHG = fabsf(DELTA);
  float FDAMP = ABSCHG / avg_blk(d.AVG, NR, g.NROWS);
  d.STORE[n] = DELTA / (1.f + FDAMP / DAMP);
}

// Corner distribution of the cell changes, mf:5828-5842, as a node-centred
// gather: node (I,J,K) receives from cells I-1..I, J..J+1, K-1..K, summed in
// the order the reference's cell loop (K, J, I) visits them.
__global__ void k_scatter(Geo g, Dev d, const float *D, float *DOUT) {
  if (gated(d)) return;
  NODE_IJK
  const size_t n = IX(I, J, K);
  float v = D[n];
  for (int kc = K - 1; kc <= K; kc++) {
    if (kc < 1 || kc > g.KMM1) continue;
    const bool top = (kc == K - 1);  // node is the cell's K+1 corner
    for (int jc = J; jc <= J + 1; jc++) {
      if (jc < 2 || jc > g.JM) continue;
      const float fac = (jc == J) ? d.FACDWN[jc - 1] : d.FACUP[jc - 1];
      for (int ic = I - 1; ic <= I; ic++) {
        if (ic < 1 || ic > g.IMM1) continue;
        const bool right = (ic == I - 1);  // node is the cell's I+1 corner
        const float *w = top ? (right ? d.FTR : d.FTL) : (right ? d.FBR : d.FBL);
        float ADD = d.STORE[IX(ic, jc, kc)];
        v = v + ADD * w[IKX(ic, kc)] * fac;
      }
    }
  }
  DOUT[n] = v;
}

// ---------------------------------------------------------------------------
// Fused TSTEP stages.  Each keeps the reference's arithmetic; only the
// launch structure differs from the phase-per-kernel form.

// mixing-plane role of J from the per-J table d.JMP (built at start-up):
// 1 = JMIX(m), 2 = JMIX(m)+1, 3 = JMIX(m)+2 for an interface m < NROWS,
// 0 elsewhere; jx returns JMIX(m)
__device__ __forceinline__ int mp_role(const Dev &d, int J, int &jx) {
  const int v = d.JMP[J - 1];
  jx = v >> 2;
  return v & 3;
}


// Residual of one (J,K) line on one wavefront: the upstream-face flux
// extrapolation at the mixing planes applied to the line's source
// (mf:5516-5555), the flux sum and SSS history with the tip-gap TFLUX
// balance (mf:5503-5510, 5560-5570), and the first pitchwise STORE average
// on the mixing-plane lines (mf:5578-5607).
// The SSS history of one conserved variable: FP32, or (the mixed form) BF16
// storage rounded to nearest even, the arithmetic in FP32 either way.
__device__ __forceinline__ unsigned short f32_to_bf16(float x) {
  unsigned u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);  // NaN stays NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(unsigned short h) { return __uint_as_float((unsigned)h << 16); }
template <bool BF>
struct Hist {
  void *p;
  __device__ __forceinline__ float ld(unsigned n) const {
    if constexpr (BF) return bf16_to_f32(((const unsigned short *)p)[n]);
    else return ((const float *)p)[n];
  }
  __device__ __forceinline__ void st(unsigned n, float v) const {
    if constexpr (BF) ((unsigned short *)p)[n] = f32_to_bf16(v);
    else ((float *)p)[n] = v;
  }
};

__global__ void k_to_bf16(const float *src, unsigned short *dst, unsigned n) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = f32_to_bf16(src[i]);
}
__global__ void k_from_bf16(const unsigned short *src, float *dst, unsigned n) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = bf16_to_f32(src[i]);
}

template <bool BF, bool BS>
__device__ __forceinline__ void resid_mp_line(const Geo &g, const Dev &d, Hist<BF> DIFF, Hist<BS> SOUT,
                                              const Params &p, int J, int K, int lane, float RATPITCH, bool tipbal,
                                              int role, int jx, float (&out)[4]) {
  const int IM = g.IM, IMM1 = g.IMM1;
  const float tavg = tipbal ? 0.5f * (d.TFLUX[IX(1, J, K)] + d.TFLUX[IX(IM, J, K)]) : 0.0f;
  const float FLOWDIRN = -d.FLOWX[IX(p.IMID, jx, K)];
  const bool fex = !(p.FEXTRAP < 0.001f);
  const bool fexline = fex && ((role == 1 && FLOWDIRN > 0.0f) || (role == 3 && !(FLOWDIRN > 0.0f)));
  float xf[4] = {0.f, 0.f, 0.f, 0.f}, AVFL = 0.0f;
  if (fexline) {
    const int jj = (role == 1) ? J - 1 : J;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int I = 64 * c + lane + 1;
      xf[c] = (I <= IMM1) ? d.XFLUX[IX(I, jj, K)] : 0.0f;
    }
    AVFL = line_sum(xf, IMM1, p.EXACT);
  }
  float st[4];
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int I = 64 * c + lane + 1;
    st[c] = 0.0f;
    if (I > IMM1) continue;
    const unsigned n = IX(I, J, K);
    float TF = d.TFLUX[n], TFP = d.TFLUX[n + 1];
    if (tipbal) {
      if (I == 1) TF = tavg;
      if (I + 1 == IM) TFP = tavg;
    }
    float src = d.SOURCE ? d.SOURCE[n] : 0.0f;  // no source array: continuity's
    if (fexline) {
      const float DFLUX = (xf[c] - AVFL * d.FP[I - 1]) * p.FEXTRAP;
      src = (role == 1) ? src - DFLUX : src + DFLUX;
    }
    if (fex && role == 2) src = 0.0f;
    const float DELTA = d.XFLUX[n] - d.XFLUX[IX(I, J - 1, K)] * RATPITCH + d.RFLUX[n] - d.RFLUX[IX(I, J, K + 1)] + TF -
                        TFP - src;
    const float df = DIFF.ld(n);
    st[c] = p.F1 * DELTA + p.F2 * df;
    DIFF.st(n, DELTA + p.F3 * df);
  }
  const bool avg1 = (role == 3 && (FLOWDIRN > 0.0f || p.FEXTRAP < 0.001f)) ||
                    (role == 1 && (FLOWDIRN < 0.0f || p.FEXTRAP < 0.001f));
  if (avg1) {
    const float SUM = line_sum(st, IMM1, p.EXACT);
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int I = 64 * c + lane + 1;
      if (I <= IMM1) st[c] = SUM * d.FP[I - 1];
    }
  }
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int I = 64 * c + lane + 1;
    if (I <= IMM1) SOUT.st(IX(I, J, K), st[c]);
    out[c] = st[c];
  }
}

// TSTEP's residual (mf:5503-5607).  One wavefront per (J, kg consecutive
// K lines), four wavefronts along J per workgroup.  Lines away from the
// mixing planes take the plain flux sum with all lines' loads in flight
// together; the mixing-plane lines go through resid_mp_line.  With part
// set, kg = KR and the wavefront's lines are one level-1 multigrid K block:
// it also leaves the fast-mode multigrid partial sums of the new STORE
// (mf:5621-5640), FP64 per level-1 I block in P[(ib, J, kb)] and per line
// group in PT[(J, kb)], lane columns summed over the block's K lines first.
#define PX(ib, J, kb) ((size_t)((ib) - 1) + (size_t)g.NIB1 * ((size_t)((J) - 1) + (size_t)g.JM * (size_t)((kb) - 1)))
template <bool BF, bool BS>
__device__ __forceinline__ void resid_line(const Geo &g, const Dev &d, Hist<BF> DIFF, Hist<BS> SOUT, const Params &p,
                                           const MgFast &m,
                                           int kg, int ir, int part, int bl) {
  __shared__ double sh[4][256];
  const int lane = threadIdx.x & 63, w = wave_id();
  const int J = (bl % gridDim.x) * 4 + w + 2, kb = bl / gridDim.x, kq = kb * kg + 1;
  if (J > g.JM) return;
  const int IM = g.IM, IMM1 = g.IMM1;
  const int nk = min(kg, g.KMM1 - kq + 1);
  // the faces, source and history of the wavefront's lines (nk of them,
  // the next line's RFLUX too) in one 64-point chunk of I: all loads of the
  // lines before any store (the arrays may alias as far as the compiler
  // knows, so a store would hold back later loads); only the wavefront's
  // own nk lines (with kg = 3 the fourth belongs to the next wavefront)
  struct Faces {
    float rf[5], xf[4], xm[4], tf[4], tp[4], sr[4], df[4];
  };
  auto load = [&](int c, Faces &F) {
    const int I = 64 * c + lane + 1;
    const bool ok = I <= IMM1;
#pragma unroll
    for (int kk = 0; kk < 4; kk++) F.xf[kk] = F.xm[kk] = F.tf[kk] = F.tp[kk] = F.sr[kk] = F.df[kk] = 0.0f;
#pragma unroll
    for (int kk = 0; kk <= 4; kk++) F.rf[kk] = (ok && kk <= nk) ? d.RFLUX[IX(I, J, min(kq + kk, g.KM))] : 0.0f;
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
      if (kk >= nk) break;
      if (!ok) continue;
      const unsigned n = IX(I, J, min(kq + kk, g.KMM1));
      F.xf[kk] = d.XFLUX[n];
      F.xm[kk] = d.XFLUX[n - IM];
      F.tf[kk] = d.TFLUX[n];
      F.tp[kk] = d.TFLUX[n + 1];
      F.sr[kk] = d.SOURCE ? d.SOURCE[n] : 0.0f;
      F.df[kk] = DIFF.ld(n);
    }
  };
  // the first chunk's loads go out before the step's gate and the row
  // lookups (their addresses depend on nothing loaded); a mixing-plane line
  // does not use them
  Faces F0;
  load(0, F0);
  if (gated(d)) return;
  const int NR = d.NROW[J - 1];
  const float RATPITCH = (float)d.NBLADE[J - 2] / (float)d.NBLADE[J - 1];
  const int kts = d.KTIPS[NR - 1], kte = d.KTIPE[NR - 1];
  int jx = 0;
  const int role = g.NROWS > 1 ? mp_role(d, J, jx) : 0;
  float st[4][4];  // [line][I chunk]: the new STORE, zero off the grid
#pragma unroll
  for (int kk = 0; kk < 4; kk++)
#pragma unroll
    for (int c = 0; c < 4; c++) st[kk][c] = 0.0f;
  if (role) {
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
      const int K = kq + kk;
      if (kk < nk) resid_mp_line(g, d, DIFF, SOUT, p, J, K, lane, RATPITCH, kts > 0 && K >= kts && K <= kte - 1, role, jx,
                                 st[kk]);
    }
  } else {
    bool tb[4];
    float tavg[4];
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
      const int K = kq + kk;
      tb[kk] = kk < nk && kts > 0 && K >= kts && K <= kte - 1;
      tavg[kk] = tb[kk] ? 0.5f * (d.TFLUX[IX(1, J, K)] + d.TFLUX[IX(IM, J, K)]) : 0.0f;
    }
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int I = 64 * c + lane + 1;
      if (64 * c >= IMM1) break;
      Faces Fc;
      if (c > 0) load(c, Fc);
      const Faces &F = c == 0 ? F0 : Fc;
      if (I > IMM1) continue;
      const float(&rf)[5] = F.rf;
      const float(&xf)[4] = F.xf, (&xm)[4] = F.xm, (&tf)[4] = F.tf, (&tp)[4] = F.tp, (&sr)[4] = F.sr,
                 (&df)[4] = F.df;
#pragma unroll
      for (int kk = 0; kk < 4; kk++) {
        if (kk >= nk) break;
        float TF = tf[kk], TFP = tp[kk];
        if (tb[kk]) {
          if (I == 1) TF = tavg[kk];
          if (I + 1 == IM) TFP = tavg[kk];
        }
        const float DELTA = xf[kk] - xm[kk] * RATPITCH + rf[kk] - rf[kk + 1] + TF - TFP - sr[kk];
        const unsigned n = IX(I, J, kq + kk);
        const float v = p.F1 * DELTA + p.F2 * df[kk];
        SOUT.st(n, v);
        DIFF.st(n, DELTA + p.F3 * df[kk]);
        st[kk][c] = v;
      }
    }
  }
  if (!part) return;
  double tot = 0.0;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const double a = (((double)st[0][c] + (double)st[1][c]) + (double)st[2][c]) + (double)st[3][c];
    sh[w][64 * c + lane] = a;
    tot += a;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off, 64);
  if (lane == 0) m.PT[JKX(J, kb + 1)] = tot;
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  for (int ib = lane + 1; ib <= g.NIB1; ib += 64) {
    const int i0 = (ib - 1) * ir, i1 = min(ib * ir, IMM1);
    double sum = 0.0;
    for (int i = i0; i < i1; i++) sum += sh[w][i];
    m.P[PX(ib, J, kb + 1)] = sum;
  }
}
// Batched over the variables of e (grid z), each with its own scratch and
// multigrid partial sums.  BF: the SSS histories in BF16; BS: the new STORE
// too (the mixed form's fused fast-mode chain).
template <bool BF, bool BS>
__device__ __forceinline__ void resid_line_eq(const Geo &g, const Dev &d0, const EqSet &e, const Params &p,
                                              const MgFast &m0, int kg, int ir, int part) {
  int z;
  const int bl = eq_block(z);
  const EqLane &q = eq_pick(e, z);
  const Dev d = eq_dev(d0, q);
  MgFast m = m0;
  m.P = q.P;
  m.PT = q.PT;
  const Hist<BF> DIFF{BF ? (void *)q.DIFF16 : (void *)q.DIFF};
  const Hist<BS> SOUT{BS ? (void *)q.STORE16 : (void *)q.STORE};
  resid_line<BF, BS>(g, d, DIFF, SOUT, p, m, kg, ir, part, bl);
}
// (the step's gate is checked inside, after the first loads)
#define RESID_LINE(NAME, BF, BS)                                                                              \
  __global__ void __launch_bounds__(256) NAME(Geo g, Dev d, EqSet e, Params p, MgFast m, int kg, int ir,       \
                                              int part) {                                                     \
    resid_line_eq<BF, BS>(g, d, e, p, m, kg, ir, part);                                                       \
  }
RESID_LINE(k_resid_line, false, false)
// the same with the SSS histories stored in BF16 (the mixed form), and with
// the new STORE in BF16 as well
RESID_LINE(k_resid_line_bf, true, false)
RESID_LINE(k_resid_line_bfs, true, true)
#undef RESID_LINE

// Fast-mode multigrid sums (mf:5621-5640), pass 2 (pass 1 is in
// k_resid_line), by block range: one thread per level-1 block (its J members from
// the level-1 list), one wavefront per level-2 block (its level-1 I and K
// blocks and J members), one workgroup per superblock (its J members, all
// level-1 K blocks); fixed-order trees throughout.
__global__ void __launch_bounds__(256) k_mg_levels(Geo g, MgFast m, EqSet e, const int *b1p, const int *b1l,
                                                   const int *sbp, const int *sbl) {
  if (m.gate && *m.gate) return;
  // grid y: the variable
  const EqLane &q = eq_pick(e, blockIdx.y);
  m.P = q.P;
  m.PT = q.PT;
  float *const B1 = q.B1CHG, *const B2 = q.B2CHG, *const SB = q.SBCHG;
  const int n1 = g.NIB1 * (g.NJB1 + 1) * g.NKB1, n2 = g.NIB2 * (g.NJB2 + 1) * g.NKB2;
  const int nblk1 = (n1 + 255) / 256, nblk2 = (n2 + 3) / 4;
  const int bx = blockIdx.x;
  if (bx < nblk1) {
    const int t = bx * 256 + threadIdx.x;
    if (t >= n1) return;
    const int nj = g.NJB1 + 1;
    const int ib = t % g.NIB1 + 1, jb = (t / g.NIB1) % nj + 1, kb = t / (g.NIB1 * nj) + 1;
    const int q0 = b1p[jb - 1], q1 = b1p[jb];
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) v[u] = (q0 + u < q1) ? m.P[PX(ib, b1l[q0 + u], kb)] : 0.0;
    double s = ((v[0] + v[1]) + v[2]) + v[3];
    for (int q = q0 + 4; q < q1; q++) s += m.P[PX(ib, b1l[q], kb)];
    B1[t] = (float)s;
    return;
  }
  if (bx < nblk1 + nblk2) {
    const int lane = threadIdx.x & 63;
    const int u = (bx - nblk1) * 4 + wave_id();
    if (u >= n2) return;
    // members: host-built P offsets of the block's (level-1 I block, J,
    // level-1 K block) partial sums
    const int q0 = m.b2_ptr[u], q1 = m.b2_ptr[u + 1];
    double s = 0.0;
    for (int q = q0 + lane; q < q1; q += 128) {
      const double a = m.P[m.b2_off[q]];
      const double b = (q + 64 < q1) ? m.P[m.b2_off[q + 64]] : 0.0;
      s += a + b;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) B2[u] = (float)s;
    return;
  }
  __shared__ double sh[256];
  const int sb = bx - nblk1 - nblk2;
  const int ja = sbp[sb], nj = sbp[sb + 1] - ja, tot = nj * g.NKB1;
  // groups of four members per thread summed in turn; four groups are
  // fetched at once (indices, then values) before they are added in order
  double s = 0.0;
  for (int c0 = 0; c0 < tot; c0 += 4096) {
    int o[16];
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int c = c0 + 256 * u + threadIdx.x;
      o[u] = (c < tot) ? (int)JKX(sbl[ja + c % nj], c / nj + 1) : -1;
    }
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; u++) v[u] = (o[u] >= 0) ? m.PT[o[u]] : 0.0;
#pragma unroll
    for (int t4 = 0; t4 < 4; t4++) s += ((v[4 * t4] + v[4 * t4 + 1]) + v[4 * t4 + 2]) + v[4 * t4 + 3];
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) SB[sb] = (float)sh[0];
}
#undef PX

// multigrid update (mf:5645-5660), one wavefront per (J, 4 consecutive K
// lines), leaving each line's FP64 sum of |STORE| for the damping average
// in red[(J,K)]
template <bool BS>
__device__ __forceinline__ void mg_apply_lines(const Geo &g, const Dev &d0, const EqSet &e) {
  const int lane = threadIdx.x & 63, w = wave_id();
  int z;
  const int bl = eq_block(z);
  const EqLane &q = eq_pick(e, z);
  const Dev d = eq_dev(d0, q);
  const Hist<BS> S{BS ? (void *)q.STORE16 : (void *)q.STORE};
  double *const red = d.RED;
  const int J = (bl % gridDim.x) * 4 + w + 2, kq = (bl / gridDim.x) * 4 + 1;
  if (J > g.JM) return;
  const int nk = min(4, g.KMM1 - kq + 1);
  // the first chunk's node values and block indices, loaded before the
  // step's gate and the J-line lookups (their addresses depend on nothing
  // loaded)
  float st0[4], sp0[4], rs0[4];
  const int I0 = min(lane + 1, g.IMM1);
  const int ib10 = d.IB1[I0 - 1], ib20 = d.IB2[I0 - 1];
#pragma unroll
  for (int kk = 0; kk < 4; kk++) {
    const unsigned n = IX(I0, J, min(kq + kk, g.KMM1));
    st0[kk] = S.ld(n);
    sp0[kk] = d.STEP[n];
    rs0[kk] = d.RSTEP[n];
  }
  if (gated(d)) return;
  const int J1 = d.JB1[J - 1], J2 = d.JB2[J - 1], JSB = d.JSBLK[J - 1];
  const float sbs = d.SBCHG[JSB - 1] * d.STEPSBK[JSB - 1];
  unsigned o1[4], o2[4];
#pragma unroll
  for (int kk = 0; kk < 4; kk++) {
    const int K = min(kq + kk, g.KMM1);
    o1[kk] = (unsigned)g.NIB1 * ((unsigned)(J1 - 1) + (unsigned)(g.NJB1 + 1) * (unsigned)(d.KB1[K - 1] - 1));
    o2[kk] = (unsigned)g.NIB2 * ((unsigned)(J2 - 1) + (unsigned)(g.NJB2 + 1) * (unsigned)(d.KB2[K - 1] - 1));
  }
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (int I = lane + 1; I <= g.IMM1; I += 64) {
    const bool first = I == lane + 1;
    const unsigned i1 = (unsigned)((first ? ib10 : d.IB1[I - 1]) - 1), i2 = (unsigned)((first ? ib20 : d.IB2[I - 1]) - 1);
    float st[4], sp[4], rs[4], c1[4], s1[4], c2[4], s2[4];
    // all loads of the four lines before any store
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
      const unsigned n = IX(I, J, min(kq + kk, g.KMM1)), b1 = i1 + o1[kk], b2 = i2 + o2[kk];
      st[kk] = first ? st0[kk] : S.ld(n);
      sp[kk] = first ? sp0[kk] : d.STEP[n];
      rs[kk] = first ? rs0[kk] : d.RSTEP[n];
      c1[kk] = d.B1CHG[b1];
      s1[kk] = d.STEP1[b1];
      c2[kk] = d.B2CHG[b2];
      s2[kk] = d.STEP2[b2];
    }
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
      if (kk >= nk) break;
      const float DELTA = st[kk] * sp[kk] + (c1[kk] * s1[kk] + c2[kk] * s2[kk] + sbs) * rs[kk];
      S.st(IX(I, J, kq + kk), DELTA);
      s[kk] += (double)fabsf(DELTA);
    }
  }
#pragma unroll
  for (int kk = 0; kk < 4; kk++) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s[kk] += __shfl_xor(s[kk], off, 64);
  }
  if (lane < nk) {
    const double v = lane == 0 ? s[0] : lane == 1 ? s[1] : lane == 2 ? s[2] : s[3];
    red[JKX(J, kq + lane)] = v;
  }
}
// (the step's gate is checked inside, after the first loads)
__global__ void __launch_bounds__(256) k_mg_apply_lines(Geo g, Dev d0, EqSet e) { mg_apply_lines<false>(g, d0, e); }
// the mixed form's fused chain: STORE in BF16
__global__ void __launch_bounds__(256) k_mg_apply_lines_bfs(Geo g, Dev d0, EqSet e) {
  mg_apply_lines<true>(g, d0, e);
}

// per-row average |change| from the (J,K) line sums (mf:5677-5693): one
// workgroup of 16 wavefronts per row (a contiguous J range); each lane sums
// its share, then fixed-order butterflies within and across wavefronts
__global__ void __launch_bounds__(1024) k_row_avg(Geo g, Dev d, EqSet e, const int *rptr, const int *rlist) {
  if (gated(d)) return;
  // grid y: the variable
  const EqLane &q = eq_pick(e, blockIdx.y);
  const double *const red = q.RED;
  __shared__ double sh[16];
  const int r = blockIdx.x, t = threadIdx.x, lane = t & 63, w = wave_id();
  const int q0 = rptr[r], nj = rptr[r + 1] - q0, ja = rlist[q0], tot = nj * g.KMM1;
  // groups of four lines per thread summed in turn, four groups fetched at
  // once before they are added in order
  double s = 0.0;
  for (int c0 = 0; c0 < tot; c0 += 16384) {
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int c = c0 + 1024 * u + t;
      v[u] = (c < tot) ? red[JKX(ja + c % nj, c / nj + 1)] : 0.0;
    }
#pragma unroll
    for (int t4 = 0; t4 < 4; t4++) s += ((v[4 * t4] + v[4 * t4 + 1]) + v[4 * t4 + 2]) + v[4 * t4 + 3];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) sh[w] = s;
  __syncthreads();
  if (w == 0) {
    double x = lane < 16 ? sh[lane] : 0.0;
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if (lane == 0) {
      const float f = (float)x;
      q.AVG[r] = f / (float)(g.IMM1 * g.KMM1 * nj);
    }
  }
}

// the damped change of one cell (mf:5716-5724)
__device__ __forceinline__ float damped(const Geo &g, const Dev &d, const float *AC, float DELTA, int J, float DAMP) {
  const float ABSCHG = fabsf(DELTA);
  const float FDAMP = ABSCHG / avg_blk(AC, d.NROW[J - 1], g.NROWS);
  return DELTA / (1.f + FDAMP / DAMP);
}

// Damping (mf:5716-5724), the second pitchwise average of the damped change
// on the mixing-plane lines (mf:5786-5818) and the corner distribution
// (mf:5828-5842).  One workgroup (4 wavefronts) per tile of kTJ x kTK node
// lines (all I).  The tile's cells (J..J+kTJ, K-1..K+kTK-1) are damped once
// into LDS and the averaged mixing-plane lines replaced by their mean there;
// the corner weights of those K planes are staged in LDS beside them.  Cells
// and weights off the grid (J < 2 or > JM, K < 1 or > KM-1, I < 1 or > IM-1)
// are stored as zeros in padded lines (column c holds cell I = c, columns 0
// and IM are pads), so each node gathers its eight corner contributions from
// cells I-1..I, J..J+1, K-1..K without a bounds test, in the order the
// reference's cell loop (K, J, I) visits them (an absent cell adds zero).
// The fast mode damps with one reciprocal per value, y / (1 + |y| / (A DAMP))
// as y * rcp(1 + |y| rAD) with rAD = 1 / (A DAMP) per J line; the exact mode
// keeps the reference's two divisions.  Dynamic LDS: kTL + kWL lines of
// `pitch` = 64 NC + 2 floats.
template <int NC, int kTK, bool BS>
__device__ __forceinline__ void damp_scatter(const Geo &g, const Dev &d, const Params &p, Hist<BS> STORE,
                                             const float *AC, const float *D, float *DOUT, int damp, int mpav,
                                             int bl) {
  constexpr int kTJ = kScatTJ, kTL = (kTJ + 1) * (kTK + 1), kMaxLn = (kTL + 3) / 4, kWL = 4 * (kTK + 1);
  static_assert(kTJ == 4, "one wavefront per tile J line");
  const float DAMP = d.SC[SC_DAMP];
  extern __shared__ float lds[];
  __shared__ float avg[kTJ + 1], rad[kTJ + 1], fdn[kTJ + 1], fup[kTJ + 1];
  __shared__ int flag[kTL];
  constexpr int pitch = 64 * NC + 2;
  const int j0 = (bl % gridDim.x) * kTJ + 1, k0 = (bl / gridDim.x) * kTK + 1;
  const int lane = threadIdx.x & 63, w = wave_id(), t = threadIdx.x;
  const int IM = g.IM, IMM1 = g.IMM1, JM = g.JM, KM = g.KM, KMM1 = g.KMM1;
  const unsigned plane = (unsigned)IM * (unsigned)JM;
  float *C = lds;                  // [kTL][pitch] cell changes, line (a, b) = J j0+a, K k0-1+b
  float *W = lds + kTL * pitch;    // [4][kTK+1][pitch]: FTL, FBL, FTR, FBR at K k0-1+b
  // this wavefront's cell lines (ln = w + 4*u) and its first chunk of
  // node values, loaded before anything waits
  float x[kMaxLn][NC];
  static_assert(NC >= 1 && NC <= 4, "IM-1 <= 256");
#pragma unroll
  for (int u = 0; u < kMaxLn; u++) {
    const int ln = w + 4 * u;
    const int jc = j0 + ln % (kTJ + 1), kc = k0 - 1 + ln / (kTJ + 1);
    const bool in = ln < kTL && jc >= 2 && jc <= JM && kc >= 1 && kc <= KMM1;
    const unsigned base = in ? (unsigned)(jc - 1) * (unsigned)IM + (unsigned)(kc - 1) * plane : 0u;
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const int I = 64 * c + lane + 1;
      x[u][c] = (in && I <= IMM1) ? STORE.ld(base + (unsigned)(I - 1)) : 0.0f;
    }
  }
  const int J = j0 + w;
  const bool jin = J <= JM;
  const unsigned nbase = jin ? (unsigned)(J - 1) * (unsigned)IM + (unsigned)(k0 - 1) * plane : 0u;
  float dv0[kTK];
#pragma unroll
  for (int kk = 0; kk < kTK; kk++) dv0[kk] = (jin && k0 + kk <= KM && lane < IM) ? D[nbase + kk * plane + lane] : 0.0f;
  // the step's gate only after these loads are out (their addresses depend
  // on nothing loaded); the whole workgroup sees the same gate
  if (gated(d)) return;
  // per-line and per-J metadata
  if (t < kTL) {
    const int jc = j0 + t % (kTJ + 1), kc = k0 - 1 + t / (kTJ + 1);
    int f = 0;
    if (jc >= 2 && jc <= JM && kc >= 1 && kc <= KMM1) {
      f = 1;
      if (mpav) {
        int jx = 0;
        const int role = mp_role(d, jc, jx);
        if (role == 1 || role == 3) {
          const float FLOWDIRN = -d.FLOWX[IX(p.IMID, jx, kc)];
          const bool on = (role == 3) ? ((FLOWDIRN > 0.0f) || (p.FEXTRAP < 0.001f))
                                      : ((FLOWDIRN < 0.0f) || (p.FEXTRAP < 0.001f));
          if (on) f = 3;
        }
      }
    }
    flag[t] = f;
  } else if (t >= 64 && t < 64 + kTJ + 1) {
    const int jc = j0 + t - 64;
    const bool jok = jc >= 2 && jc <= JM;
    const float A = (jok && damp) ? avg_blk(AC, d.NROW[jc - 1], g.NROWS) : 1.0f;
    avg[t - 64] = A;
    rad[t - 64] = 1.0f / (A * DAMP);
    fdn[t - 64] = jok ? d.FACDWN[jc - 1] : 0.0f;
    fup[t - 64] = jok ? d.FACUP[jc - 1] : 0.0f;
  }
  // corner weights: 4 arrays x (kTK+1) K planes, kWL / 4 lines per
  // wavefront, zero off the grid and in the pad columns
#pragma unroll
  for (int u = 0; u < kWL / 4; u++) {
    const int wl = w + 4 * u, arr = wl / (kTK + 1), kc = k0 - 1 + wl % (kTK + 1);
    // wavefront-uniform selects (an indexed table of the four pointers
    // would be built in scratch memory, in every lane)
    const float *src = d.FTL;
    if (arr == 1) src = d.FBL;
    if (arr == 2) src = d.FTR;
    if (arr == 3) src = d.FBR;
    const bool kin = kc >= 1 && kc <= KMM1;
    const unsigned kb = kin ? (unsigned)(kc - 1) * (unsigned)IM : 0u;
    float *Wl = W + wl * pitch;
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const int I = 64 * c + lane + 1;
      Wl[I] = (kin && I <= IMM1) ? src[kb + (unsigned)(I - 1)] : 0.0f;
    }
    if (lane == 0) {
      Wl[0] = 0.0f;
      Wl[64 * NC + 1] = 0.0f;
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kMaxLn; u++) {
    const int ln = w + 4 * u;
    if (ln >= kTL) continue;
    const int f = flag[ln];
    float v[4] = {0.f, 0.f, 0.f, 0.f};
    if (f) {
      const int a = ln % (kTJ + 1);
#pragma unroll
      for (int c = 0; c < NC; c++) {
        float y = x[u][c];
        if (damp) {
          if (p.EXACT) {
            const float FDAMP = fabsf(y) / avg[a];
            y = y / (1.f + FDAMP / DAMP);
          } else {
            y = y * __builtin_amdgcn_rcpf(1.f + fabsf(y) * rad[a]);
          }
        }
        v[c] = y;
      }
      if (f == 3) {
        const float SUM = line_sum(v, IMM1, p.EXACT) / (float)IMM1;
#pragma unroll
        for (int c = 0; c < NC; c++) {
          const int I = 64 * c + lane + 1;
          v[c] = I <= IMM1 ? SUM : 0.0f;
        }
      }
    }
    float *Cl = C + ln * pitch;
#pragma unroll
    for (int c = 0; c < NC; c++) Cl[64 * c + lane + 1] = v[c];
    if (lane == 0) {
      Cl[0] = 0.0f;
      Cl[64 * NC + 1] = 0.0f;
    }
  }
  __syncthreads();
  // gather: wavefront w takes node lines J = j0 + w, K = k0 .. k0+kTK-1
  if (!jin) return;
  const int nk = min(kTK, KM - k0 + 1);
  const float fd0 = fdn[w], fu1 = fup[w + 1];
  // node chunks: IM may be 64 NC + 1 (the last node's cells are I-1 and the
  // right-hand pad)
#pragma unroll
  for (int c = 0; c <= NC; c++) {
    const int I = 64 * c + lane + 1;
    if (64 * c >= IM) break;
    float dv[kTK];
#pragma unroll
    for (int kk = 0; kk < kTK; kk++)
      dv[kk] = c == 0 ? dv0[kk] : (kk < nk && I <= IM) ? D[nbase + kk * plane + (unsigned)(I - 1)] : 0.0f;
#pragma unroll
    for (int kk = 0; kk < kTK; kk++) {
      if (kk >= nk) break;
      float v = dv[kk];
#pragma unroll
      for (int b2 = 0; b2 < 2; b2++) {
        const int b = kk + b2;        // cell K line kc = k0 - 1 + b = K - 1 + b2
        const bool top = (b2 == 0);  // node is the cell's K+1 corner
        // weights of the cells I-1 (node is their I+1 corner) and I
        const float wr = W[((top ? 2 : 3) * (kTK + 1) + b) * pitch + I - 1];
        const float wl = W[((top ? 0 : 1) * (kTK + 1) + b) * pitch + I];
#pragma unroll
        for (int a2 = 0; a2 < 2; a2++) {
          const float fac = a2 == 0 ? fd0 : fu1;
          const float *Lc = C + (w + a2 + (kTJ + 1) * b) * pitch;
          v = v + Lc[I - 1] * wr * fac;
          v = v + Lc[I] * wl * fac;
        }
      }
      if (I <= IM) DOUT[nbase + kk * plane + (unsigned)(I - 1)] = v;
    }
  }
}
// batched over the variables of e (grid z): D = q.D, the result to q.DSC
// (with BS, the mixed form's fused chain, STORE read from its BF16 copy);
// the step's gate is checked inside, after the first loads
#define DAMP_SCATTER_BS(NC, TK, BS, SFX)                                                                     \
  __global__ void __launch_bounds__(256) k_damp_scatter##NC##_k##TK##SFX(Geo g, Dev d0, EqSet e, Params p,    \
                                                                         int damp, int mpav) {               \
    int z;                                                                                                   \
    const int bl = eq_block(z);                                                                              \
    const EqLane &q = eq_pick(e, z);                                                                         \
    const Hist<BS> S{BS ? (void *)q.STORE16 : (void *)q.STORE};                                             \
    damp_scatter<NC, TK, BS>(g, d0, p, S, q.AVG, q.D, q.DSC, damp, mpav, bl);                                \
  }
#define DAMP_SCATTER(NC, TK) DAMP_SCATTER_BS(NC, TK, false, ) DAMP_SCATTER_BS(NC, TK, true, _bfs)
DAMP_SCATTER(1, 2)
DAMP_SCATTER(1, 4)
DAMP_SCATTER(1, 8)
DAMP_SCATTER(2, 2)
DAMP_SCATTER(2, 4)
DAMP_SCATTER(2, 8)
DAMP_SCATTER(4, 2)
DAMP_SCATTER(4, 4)
DAMP_SCATTER(4, 8)
#undef DAMP_SCATTER
#undef DAMP_SCATTER_BS

// SMOOTH_VAR (mf:15352-15594) followed by TSTEP's periodic, tip-gap and
// mixing-plane conditions on D (mf:5858-5926).  One workgroup (8
// wavefronts) per J plane: wavefront w takes the K lines w+1, w+9, ...,
// lane l the I points l+1, l+65, ...  The streamwise stencil reads the
// pre-smoothing field SRC (global memory, L2-resident) four K lines at a
// time with all loads in flight; the rest of the work is plane-local and
// runs in LDS, with the pitchwise and spanwise coefficients staged there.
// Writes the finished plane to DST.  Dynamic LDS: 3*IM*KM + 2*(IM+KM)
// floats.
__device__ __forceinline__ void smooth_plane(const Geo &g, const Dev &d, const EqLane &q, const Params &p,
                                             const StepScalars &s, int J, float *lds) {
  const float *const SRC = q.DSC;
  float *const DST = q.D;
  const int ncall = q.ncall;
  const int lane = threadIdx.x & 63, w = wave_id();
  const int IM = g.IM, KM = g.KM, IMM1 = g.IMM1, IMM2 = IM - 2, KMM1 = g.KMM1, KMM2 = KM - 2;
  const int npl = IM * KM;
  float *PL = lds;            // the plane, [K][I]
  float *AV = lds + npl;      // AVG scratch
  float *CU = lds + 2 * npl;  // CURVE scratch
  float *cFD = lds + 3 * npl, *cFU = cFD + IM, *cFKD = cFU + IM, *cFKU = cFKD + KM;
#define PLX(I, K) PL[((I) - 1) + IM * ((K) - 1)]
#define AVX(I, K) AV[((I) - 1) + IM * ((K) - 1)]
#define CUX(I, K) CU[((I) - 1) + IM * ((K) - 1)]
#define FOR_PLANE for (int K = w + 1; K <= KM; K += 8) for (int I = lane + 1; I <= IM; I += 64)
  for (int t = threadIdx.x; t < IM; t += 512) {
    cFD[t] = d.FD[t];
    cFU[t] = d.FU[t];
  }
  for (int t = threadIdx.x; t < KM; t += 512) {
    cFKD[t] = d.FKD[t];
    cFKU[t] = d.FKU[t];
  }
  const float SFX = d.SC[SC_SFX], SFX1 = 1.f - SFX, FAC_4TH = p.FAC_4TH;
  const float SFXB = p.FSMTHB * SFX, SFXB1 = 1.0f - SFXB, SFXBH = 0.5f * SFXB;
  const float SFXH = 0.5f * SFX, SFXHM1 = 1.f - SFXH, SFX14 = 0.25f * SFX;
  const float SFT = d.SC[SC_SFT], SFT1 = 1.0f - SFT, SFTH = 0.5f * SFT;
  const int NR = d.NROW[J - 1];
  const int J1 = d.JSTART[NR - 1], J2 = d.JMIX[NR - 1], JLED = d.JLED[NR - 1];
  // streamwise smoothing, mf:15369-15445.  The case depends on J only.
  const bool smx = !(SFX < 0.0001f);
  int cs = 0;  // 0 copy, 1 interior, 2 J1+2, 3 J1+1, 4 J2-2, 5 J2-1
  if (smx) {
    if (J >= J1 + 3 && J <= J2 - 3) cs = 1;
    if (J == J1 + 2) cs = 2;
    if (J == J1 + 1) cs = 3;
    if (J == J2 - 2) cs = 4;
    if (J == J2 - 1) cs = 5;
  }
  // four K lines per wavefront per pass; K is clamped so every load is
  // unconditional and the loads of all four lines issue together
  auto stream = [&](auto f) {
    for (int K0 = w + 1; K0 <= KM; K0 += 32)
      for (int I = lane + 1; I <= IM; I += 64) {
        float v[4];
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = f(I, min(K0 + 8 * q, KM));
#pragma unroll
        for (int q = 0; q < 4; q++)
          if (K0 + 8 * q <= KM) PLX(I, K0 + 8 * q) = v[q];
      }
  };
#define S(JJ) SRC[IX(I, (JJ), K)]
  switch (cs) {
    case 1:
      stream([&](int I, int K) {
        const float AVERG = 0.5f * (S(J + 1) + S(J - 1));
        const float CURV = AVERG - 0.5f * S(J) - 0.25f * (S(J - 2) + S(J + 2));
        return SFX1 * S(J) + SFX * (AVERG + FAC_4TH * CURV);
      });
      break;
    case 2: stream([&](int I, int K) { return SFXB1 * S(J1 + 2) + SFXBH * (S(J1 + 3) + S(J1 + 1)); }); break;
    case 3:
      stream([&](int I, int K) {
        return SFXB1 * S(J1 + 1) + SFXB * (1.25f * S(J1 + 2) + 0.5f * S(J1 + 3) - 0.75f * S(J1 + 4));
      });
      break;
    case 4: stream([&](int I, int K) { return SFXB1 * S(J2 - 2) + SFXBH * (S(J2 - 1) + S(J2 - 3)); }); break;
    case 5:
      stream([&](int I, int K) {
        return SFXB1 * S(J2 - 1) + SFXB * (1.25f * S(J2 - 2) + 0.5f * S(J2 - 3) - 0.75f * S(J2 - 4));
      });
      break;
    default: stream([&](int I, int K) { return S(J); }); break;
  }
  // leading-edge points I = 1 and IM (mf:15417-15434), after the above
  if (smx && (J == JLED - 2 || J == JLED - 1 || J == JLED)) {
    __syncthreads();
    for (int t = threadIdx.x; t < 2 * KM; t += blockDim.x) {
      const int K = t % KM + 1, I = (t < KM) ? 1 : IM, Io = (I == 1) ? IM : 1;
      PLX(I, K) = (J == JLED) ? SFXHM1 * S(J) + SFX14 * (SRC[IX(I, J + 1, K)] + SRC[IX(Io, J, K)])
                              : SFXHM1 * S(J) + SFX14 * (SRC[IX(2, J, K)] + SRC[IX(IMM1, J, K)]);
    }
  }
#undef S
  __syncthreads();
  // pitchwise and spanwise smoothing and corners, mf:15460-15541 (J >= 2)
  if (J >= 2 && !(SFT < 0.0001f)) {
    const bool blade = (d.IND[J - 1] == 1 || d.INDLE[J - 1] == 1);
    FOR_PLANE {
      float avg, cur;
      if (I >= 2 && I <= IMM1) {
        avg = cFD[I - 1] * PLX(I + 1, K) + cFU[I - 1] * PLX(I - 1, K);
        cur = PLX(I, K) - avg;
      } else if (!blade) {
        avg = 0.5f * (PLX(2, K) + PLX(IMM1, K));
        cur = PLX(1, K) - avg;
      } else {
        avg = (I == 1) ? PLX(2, K) + cFU[0] * (PLX(2, K) - PLX(3, K))
                       : PLX(IMM1, K) + cFU[IM - 1] * (PLX(IMM1, K) - PLX(IMM2, K));
        cur = 0.0f;
      }
      AVX(I, K) = avg;
      CUX(I, K) = cur;
    }
    __syncthreads();
    FOR_PLANE {
      float sc;
      if (I >= 2 && I <= IMM1) sc = cFU[I - 1] * CUX(I - 1, K) + cFD[I - 1] * CUX(I + 1, K);
      else if (!blade) sc = 0.5f * (CUX(2, K) + CUX(IMM1, K));
      else sc = 0.0f;
      // AVG(I,K) is read only by this thread, so the result can replace it
      AVX(I, K) = SFT1 * PLX(I, K) + SFT * (AVX(I, K) + FAC_4TH * sc);
    }
    __syncthreads();
    { float *t = PL; PL = AV; AV = t; }
    // spanwise
    FOR_PLANE {
      float avg, cur;
      if (K >= 2 && K <= KMM1) {
        avg = cFKD[K - 1] * PLX(I, K + 1) + cFKU[K - 1] * PLX(I, K - 1);
        cur = PLX(I, K) - avg;
      } else if (K == 1) {
        avg = PLX(I, 2) + 0.5f * cFKU[0] * (PLX(I, 2) - PLX(I, 3));
        cur = 0.0f;
      } else {
        avg = PLX(I, KMM1) + 0.5f * cFKU[KM - 1] * (PLX(I, KMM1) - PLX(I, KMM2));
        cur = 0.0f;
      }
      AVX(I, K) = avg;
      CUX(I, K) = cur;
    }
    __syncthreads();
    FOR_PLANE {
      const float sc = (K >= 2 && K <= KMM1) ? cFKU[K - 1] * CUX(I, K - 1) + cFKD[K - 1] * CUX(I, K + 1) : 0.0f;
      AVX(I, K) = SFT1 * PLX(I, K) + SFT * (AVX(I, K) + FAC_4TH * sc);
    }
    __syncthreads();
    { float *t = PL; PL = AV; AV = t; }
    if (threadIdx.x == 0) {
      PLX(1, 1) = SFT1 * PLX(1, 1) + SFTH * (PLX(2, 1) + PLX(1, 2));
      PLX(IM, 1) = SFT1 * PLX(IM, 1) + SFTH * (PLX(IMM1, 1) + PLX(IM, 2));
      PLX(1, KM) = SFT1 * PLX(1, KM) + SFTH * (PLX(1, KMM1) + PLX(2, KM));
      PLX(IM, KM) = SFT1 * PLX(IM, KM) + SFTH * (PLX(IMM1, KM) + PLX(IM, KMM1));
    }
    __syncthreads();
  }
  // exit smoothing, mf:15553-15588: Gauss-Seidel sweeps in I, then in K
  if (!(p.SFEXIT < 0.0001f) && J >= g.JM + 1 - p.NSFEXIT) {
    const float SFEX1 = s.SFEX1, SFEXH = s.SFEXH, SFEXIT = p.SFEXIT;
    // (the running value is carried in a register, so a point's update
    // waits on the previous one's arithmetic and not on an LDS round trip)
    for (int K = threadIdx.x + 1; K <= KM; K += blockDim.x) {
      float prev = PLX(1, K), cur = PLX(2, K);
      for (int I = 2; I <= IMM1; I++) {
        const float nxt = PLX(I + 1, K);
        const float v = SFEX1 * cur + SFEXH * (nxt + prev);
        PLX(I, K) = v;
        prev = v;
        cur = nxt;
      }
      PLX(1, K) = PLX(1, K) + SFEXIT * (2.f * PLX(2, K) - PLX(3, K) - PLX(1, K));
      PLX(IM, K) = PLX(IM, K) + SFEXIT * (2.f * PLX(IMM1, K) - PLX(IMM2, K) - PLX(IM, K));
    }
    __syncthreads();
    for (int I = threadIdx.x + 1; I <= IM; I += blockDim.x) {
      float prev = PLX(I, 1), cur = PLX(I, 2);
      for (int K = 2; K <= KMM1; K++) {
        const float nxt = PLX(I, K + 1);
        const float v = SFEX1 * cur + SFEXH * (prev + nxt);
        PLX(I, K) = v;
        prev = v;
        cur = nxt;
      }
      PLX(I, 1) = PLX(I, 1) + SFEXIT * (2.f * PLX(I, 2) - PLX(I, 3) - PLX(I, 1));
      PLX(I, KM) = PLX(I, KM) + SFEXIT * (2.f * PLX(I, KMM1) - PLX(I, KMM2) - PLX(I, KM));
    }
    __syncthreads();
  }
  // periodicity off the blade and across the tip gap, mf:5858-5891
  const int ILAST = (ncall == 2) ? IMM1 : IM;
  {
    const bool per = !((d.IND[J - 1] == 1) || (d.INDLE[J - 1] == 1));
    int K1 = 1, K2 = 0;
    if (d.KTIPS[NR - 1] > 0) {
      K1 = d.KTIPS[NR - 1];
      K2 = d.KTIPE[NR - 1];
      if (K1 == 1) K2 = K2 - 1;
      if (K2 == KM) K1 = K1 + 1;
    }
    for (int K = threadIdx.x + 1; K <= KM; K += blockDim.x) {
      if (per) {
        PLX(1, K) = 0.5f * (PLX(1, K) + PLX(ILAST, K));
        PLX(ILAST, K) = PLX(1, K);
      }
      if (K >= K1 && K <= K2) {
        PLX(1, K) = 0.5f * (PLX(1, K) + PLX(ILAST, K));
        PLX(ILAST, K) = PLX(1, K);
      }
    }
  }
  __syncthreads();
  // pitchwise-uniform D on both faces of a mixing plane, mf:5901-5926
  if (p.IFMIX != 0) {
    bool mp = (J == J2 && NR < g.NROWS) || (NR > 1 && J == J1);
    if (mp) {
      // one wavefront per K line: the terms in parallel, summed in the
      // reference's order (EXACT) or by the FP64 butterfly
      for (int K = w + 1; K <= KM; K += 8) {
        float v[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
          const int I = 64 * c + lane + 1;
          v[c] = (I <= IMM1) ? d.FP[I - 1] * (PLX(I, K) + PLX(I + 1, K)) : 0.0f;
        }
        const float AVG = 0.5f * line_sum(v, IMM1, p.EXACT);
        for (int I = lane + 1; I <= IM; I += 64) PLX(I, K) = AVG;
      }
      __syncthreads();
    }
  }
  FOR_PLANE DST[IX(I, J, K)] = PLX(I, K);
#undef FOR_PLANE
#undef PLX
#undef AVX
#undef CUX
}
__global__ void __launch_bounds__(512, 6) k_smooth_plane(Geo g, Dev d, EqSet e, Params p, StepScalars s, int j0) {
  if (gated(d)) return;
  extern __shared__ float lds[];
  // batched over the variables of e (grid z): q.DSC smoothed into q.D
  int z;
  const int J = eq_block(z) + j0;
  smooth_plane(g, d, eq_pick(e, z), p, s, J, lds);
}

// SMOOTH_VAR and TSTEP's D conditions as k_smooth_plane, for IM <= 64 and
// planes without exit smoothing, with the plane held in registers: wavefront
// w owns K lines w*L+1 .. w*L+L, lane l the point I = l+1.  The pitchwise
// stencil and the line conditions use lane moves (DPP shifts to the
// neighbours, readlane for the fixed lanes 1, 2, 3, IM-2, IM-1 and the last
// point); only the spanwise stencil goes through LDS (the whole plane, one
// barrier), and the corner update reads lines 2 and KM-1 from LDS.  Offsets
// are 32-bit (the grid has < 2^31 nodes).
__device__ __forceinline__ float lane_next(float x) {  // lane l gets lane l+1 (lane 63: itself)
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(x), __float_as_int(x), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_prev(float x) {  // lane l gets lane l-1 (lane 0: itself)
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(x), __float_as_int(x), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_at(float x, int l) {  // lane l's value, wavefront-uniform l
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

// LDS floats of smooth_reg<L, NW, NC>: the plane, then lines 2 and KM-1
template <int L, int NW, int NC>
constexpr int smooth_reg_lds() { return L * NW * 64 * NC + 2 * 64 * NC; }

template <int L, int NW, int NC>
__device__ __forceinline__ void smooth_reg(const Geo &g, const Dev &d, const float *SRC, float *DST, const Params &p,
                                           const StepScalars &s, int ncall, int J, float *lds) {
  // NC chunks of 64 points per lane: point I = 64 c + lane + 1 of chunk c;
  // the NW wavefronts' L lines each cover the plane's KM <= L NW lines
  float *const PLs = lds;
  float *const C2r = lds + L * NW * 64 * NC;
  auto C2 = [&](int which) { return C2r + which * 64 * NC; };
  constexpr int P = 64 * NC;  // LDS line pitch
  const int lane = threadIdx.x & 63, w = wave_id();
  const int IM = g.IM, KM = g.KM, IMM1 = g.IMM1, IMM2 = IM - 2, KMM1 = g.KMM1, KMM2 = KM - 2;
  int Iv[NC], Icv[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) {
    Iv[c] = 64 * c + lane + 1;
    Icv[c] = min(Iv[c], IM);
  }
  // the point I+1, I-1 of chunk c (across the chunk boundary by readlane)
  auto nxt = [&](const float (&x)[NC], int c) -> float {
    float r = lane_next(x[c]);
    if (c + 1 < NC) {
      const float h = lane_at(x[c + 1 < NC ? c + 1 : c], 0);
      r = (lane == 63) ? h : r;
    }
    return r;
  };
  auto prv = [&](const float (&x)[NC], int c) -> float {
    float r = lane_prev(x[c]);
    if (c > 0) {
      const float h = lane_at(x[c > 0 ? c - 1 : 0], 63);
      r = (lane == 0) ? h : r;
    }
    return r;
  };
  // the value of point i + 1 (i wavefront-uniform)
  auto at = [&](const float (&x)[NC], int i) -> float {
    const int c = i >> 6, l = i & 63;
    float r = lane_at(x[0], l);
#pragma unroll
    for (int q = 1; q < NC; q++)
      if (c == q) r = lane_at(x[q], l);
    return r;
  };
  const int kbase = w * L + 1;
  const unsigned plane = (unsigned)IM * (unsigned)g.JM;
  const float SFX = d.SC[SC_SFX], SFX1 = 1.f - SFX, FAC_4TH = p.FAC_4TH;
  const float SFXB = p.FSMTHB * SFX, SFXB1 = 1.0f - SFXB, SFXBH = 0.5f * SFXB;
  const float SFXH = 0.5f * SFX, SFXHM1 = 1.f - SFXH, SFX14 = 0.25f * SFX;
  const float SFT = d.SC[SC_SFT], SFT1 = 1.0f - SFT, SFTH = 0.5f * SFT;
  const int NR = d.NROW[J - 1];
  const int J1 = d.JSTART[NR - 1], J2 = d.JMIX[NR - 1], JLED = d.JLED[NR - 1];
  const bool smx = !(SFX < 0.0001f);
  int cs = 0;  // 0 copy, 1 interior, 2 J1+2, 3 J1+1, 4 J2-2, 5 J2-1
  if (smx) {
    if (J >= J1 + 3 && J <= J2 - 3) cs = 1;
    if (J == J1 + 2) cs = 2;
    if (J == J1 + 1) cs = 3;
    if (J == J2 - 2) cs = 4;
    if (J == J2 - 1) cs = 5;
  }
  // the spanwise coefficients of this wavefront's lines and their
  // neighbours (FKU, FKD at K-2 .. K+L), read before the plane loads
  float fku[L + 2], fkd[L + 2];
#pragma unroll
  for (int l = 0; l < L + 2; l++) {
    const int K = min(max(kbase + l - 1, 1), KM);
    fku[l] = d.FKU[K - 1];
    fkd[l] = d.FKD[K - 1];
  }
  // streamwise smoothing, mf:15369-15445 (loads clamped into the grid)
  float v[L][NC];
  auto stream = [&](auto f) {
#pragma unroll
    for (int l = 0; l < L; l++)
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const unsigned o = (unsigned)(min(kbase + l, KM) - 1) * plane + (unsigned)(Icv[c] - 1);
        v[l][c] = f(o);
      }
  };
#define S(JJ) SRC[o + (unsigned)((JJ) - 1) * (unsigned)IM]
  switch (cs) {
    case 1:
      stream([&](unsigned o) {
        const float AVERG = 0.5f * (S(J + 1) + S(J - 1));
        const float CURV = AVERG - 0.5f * S(J) - 0.25f * (S(J - 2) + S(J + 2));
        return SFX1 * S(J) + SFX * (AVERG + FAC_4TH * CURV);
      });
      break;
    case 2: stream([&](unsigned o) { return SFXB1 * S(J1 + 2) + SFXBH * (S(J1 + 3) + S(J1 + 1)); }); break;
    case 3:
      stream([&](unsigned o) {
        return SFXB1 * S(J1 + 1) + SFXB * (1.25f * S(J1 + 2) + 0.5f * S(J1 + 3) - 0.75f * S(J1 + 4));
      });
      break;
    case 4: stream([&](unsigned o) { return SFXB1 * S(J2 - 2) + SFXBH * (S(J2 - 1) + S(J2 - 3)); }); break;
    case 5:
      stream([&](unsigned o) {
        return SFXB1 * S(J2 - 1) + SFXB * (1.25f * S(J2 - 2) + 0.5f * S(J2 - 3) - 0.75f * S(J2 - 4));
      });
      break;
    default: stream([&](unsigned o) { return S(J); }); break;
  }
  // leading-edge points I = 1 and IM (mf:15417-15434)
  if (smx && (J == JLED - 2 || J == JLED - 1 || J == JLED)) {
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const int I = Iv[c];
      if (I != 1 && I != IM) continue;
      const int Io = (I == 1) ? IM : 1;
#pragma unroll
      for (int l = 0; l < L; l++) {
        const int K = min(kbase + l, KM);
        const unsigned o = (unsigned)(K - 1) * plane + (unsigned)(Icv[c] - 1);
        v[l][c] = (J == JLED) ? SFXHM1 * S(J) + SFX14 * (SRC[IX(I, J + 1, K)] + SRC[IX(Io, J, K)])
                              : SFXHM1 * S(J) + SFX14 * (SRC[IX(2, J, K)] + SRC[IX(IMM1, J, K)]);
      }
    }
  }
#undef S
  // pitchwise and spanwise smoothing and corners, mf:15460-15541 (J >= 2)
  if (J >= 2 && !(SFT < 0.0001f)) {
    const bool blade = (d.IND[J - 1] == 1 || d.INDLE[J - 1] == 1);
    const float FU1 = d.FU[0], FUM = d.FU[IM - 1];
    float FDI[NC], FUI[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
      FDI[c] = d.FD[Icv[c] - 1];
      FUI[c] = d.FU[Icv[c] - 1];
    }
#pragma unroll
    for (int l = 0; l < L; l++) {
      const float x1 = at(v[l], 0), x2 = at(v[l], 1), x3 = at(v[l], 2);
      const float xm1 = at(v[l], IMM1 - 1), xm2 = at(v[l], IMM2 - 1);
      float avg[NC], cur[NC];
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const int I = Iv[c];
        const float x = v[l][c];
        const float xr = nxt(v[l], c), xl = prv(v[l], c);
        if (I >= 2 && I <= IMM1) {
          avg[c] = FDI[c] * xr + FUI[c] * xl;
          cur[c] = x - avg[c];
        } else if (!blade) {
          avg[c] = 0.5f * (x2 + xm1);
          cur[c] = x1 - avg[c];
        } else {
          avg[c] = (I == 1) ? x2 + FU1 * (x2 - x3) : xm1 + FUM * (xm1 - xm2);
          cur[c] = 0.0f;
        }
      }
      const float c2 = at(cur, 1), cm1 = at(cur, IMM1 - 1);
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const int I = Iv[c];
        const float cr = nxt(cur, c), cl = prv(cur, c);
        float sc;
        if (I >= 2 && I <= IMM1) sc = FUI[c] * cl + FDI[c] * cr;
        else if (!blade) sc = 0.5f * (c2 + cm1);
        else sc = 0.0f;
        v[l][c] = SFT1 * v[l][c] + SFT * (avg[c] + FAC_4TH * sc);
      }
    }
    // spanwise, from the whole pitchwise-smoothed plane in LDS
#pragma unroll
    for (int l = 0; l < L; l++)
#pragma unroll
      for (int c = 0; c < NC; c++)
        if (kbase + l <= KM && Iv[c] <= IM) PLs[(kbase + l - 1) * P + 64 * c + lane] = v[l][c];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NC; c++) {
#define PK(KK) PLs[((KK) - 1) * P + 64 * c + lane]
#pragma unroll
      for (int l = 0; l < L; l++) {
        const int K = kbase + l;
        if (K > KM) break;
        const float x = v[l][c];
        float avg;
        float sc = 0.0f;
        // coefficients at K-1, K, K+1: fk?[l], fk?[l+1], fk?[l+2]
        if (K >= 2 && K <= KMM1) {
          avg = fkd[l + 1] * PK(K + 1) + fku[l + 1] * PK(K - 1);
          // CURVE at K-1 and K+1 (zero on the end lines)
          float cm = 0.0f, cp = 0.0f;
          if (K - 1 >= 2) cm = PK(K - 1) - (fkd[l] * PK(K) + fku[l] * PK(K - 2));
          if (K + 1 <= KMM1) cp = PK(K + 1) - (fkd[l + 2] * PK(K + 2) + fku[l + 2] * PK(K));
          sc = fku[l + 1] * cm + fkd[l + 1] * cp;
        } else if (K == 1) {
          avg = PK(2) + 0.5f * fku[l + 1] * (PK(2) - PK(3));
        } else {
          avg = PK(KMM1) + 0.5f * fku[l + 1] * (PK(KMM1) - PK(KMM2));
        }
        v[l][c] = SFT1 * x + SFT * (avg + FAC_4TH * sc);
      }
#undef PK
    }
    // corners: lines 2 and KM-1 to LDS, then the owners of lines 1 and KM
#pragma unroll
    for (int l = 0; l < L; l++) {
      const int K = kbase + l;
#pragma unroll
      for (int c = 0; c < NC; c++) {
        if (K == 2) C2(0)[64 * c + lane] = v[l][c];
        if (K == KMM1) C2(1)[64 * c + lane] = v[l][c];
      }
    }
    __syncthreads();
#pragma unroll
    for (int l = 0; l < L; l++) {
      const int K = kbase + l;
      if (K != 1 && K != KM) continue;
      float xn[NC], xp[NC];
#pragma unroll
      for (int c = 0; c < NC; c++) {
        xn[c] = nxt(v[l], c);
        xp[c] = prv(v[l], c);
      }
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const int I = Iv[c];
        const float x = v[l][c];
        const float y = (K == 1) ? C2(0)[64 * c + lane] : C2(1)[64 * c + lane];
        if (K == 1) {
          if (I == 1) v[l][c] = SFT1 * x + SFTH * (xn[c] + y);
          if (I == IM) v[l][c] = SFT1 * x + SFTH * (xp[c] + y);
        } else {
          if (I == 1) v[l][c] = SFT1 * x + SFTH * (y + xn[c]);
          if (I == IM) v[l][c] = SFT1 * x + SFTH * (xp[c] + y);
        }
      }
    }
  }
  // exit smoothing, mf:15553-15588: the plane through LDS for the
  // Gauss-Seidel sweeps in I (a thread per K line), then in K (a thread per
  // I line), the running value carried in a register; the pitchwise and
  // spanwise stages are done with PLs (the corner stage's barrier)
  if (!(p.SFEXIT < 0.0001f) && J >= g.JM + 1 - p.NSFEXIT) {
    const float SFEX1 = s.SFEX1, SFEXH = s.SFEXH, SFEXIT = p.SFEXIT;
#define PX(I, K) PLs[((K) - 1) * P + (I) - 1]
#pragma unroll
    for (int l = 0; l < L; l++)
#pragma unroll
      for (int c = 0; c < NC; c++)
        if (kbase + l <= KM && Iv[c] <= IM) PX(Iv[c], kbase + l) = v[l][c];
    __syncthreads();
    for (int K = threadIdx.x + 1; K <= KM; K += blockDim.x) {
      float prev = PX(1, K), cur = PX(2, K);
      for (int I = 2; I <= IMM1; I++) {
        const float nxt = PX(I + 1, K);
        const float x = SFEX1 * cur + SFEXH * (nxt + prev);
        PX(I, K) = x;
        prev = x;
        cur = nxt;
      }
      PX(1, K) = PX(1, K) + SFEXIT * (2.f * PX(2, K) - PX(3, K) - PX(1, K));
      PX(IM, K) = PX(IM, K) + SFEXIT * (2.f * PX(IMM1, K) - PX(IMM2, K) - PX(IM, K));
    }
    __syncthreads();
    for (int I = threadIdx.x + 1; I <= IM; I += blockDim.x) {
      float prev = PX(I, 1), cur = PX(I, 2);
      for (int K = 2; K <= KMM1; K++) {
        const float nxt = PX(I, K + 1);
        const float x = SFEX1 * cur + SFEXH * (prev + nxt);
        PX(I, K) = x;
        prev = x;
        cur = nxt;
      }
      PX(I, 1) = PX(I, 1) + SFEXIT * (2.f * PX(I, 2) - PX(I, 3) - PX(I, 1));
      PX(I, KM) = PX(I, KM) + SFEXIT * (2.f * PX(I, KMM1) - PX(I, KMM2) - PX(I, KM));
    }
    __syncthreads();
#pragma unroll
    for (int l = 0; l < L; l++)
#pragma unroll
      for (int c = 0; c < NC; c++)
        if (kbase + l <= KM && Iv[c] <= IM) v[l][c] = PX(Iv[c], kbase + l);
#undef PX
  }
  // periodicity off the blade and across the tip gap, mf:5858-5891
  const int ILAST = (ncall == 2) ? IMM1 : IM;
  const bool per = !((d.IND[J - 1] == 1) || (d.INDLE[J - 1] == 1));
  int K1 = 1, K2 = 0;
  if (d.KTIPS[NR - 1] > 0) {
    K1 = d.KTIPS[NR - 1];
    K2 = d.KTIPE[NR - 1];
    if (K1 == 1) K2 = K2 - 1;
    if (K2 == KM) K1 = K1 + 1;
  }
#pragma unroll
  for (int l = 0; l < L; l++) {
    const int K = kbase + l;
    if (per) {
      const float a = 0.5f * (at(v[l], 0) + at(v[l], ILAST - 1));
#pragma unroll
      for (int c = 0; c < NC; c++)
        if (Iv[c] == 1 || Iv[c] == ILAST) v[l][c] = a;
    }
    if (K >= K1 && K <= K2) {
      const float a = 0.5f * (at(v[l], 0) + at(v[l], ILAST - 1));
#pragma unroll
      for (int c = 0; c < NC; c++)
        if (Iv[c] == 1 || Iv[c] == ILAST) v[l][c] = a;
    }
  }
  // pitchwise-uniform D on both faces of a mixing plane, mf:5901-5926
  if (p.IFMIX != 0 && ((J == J2 && NR < g.NROWS) || (NR > 1 && J == J1))) {
#pragma unroll
    for (int l = 0; l < L; l++) {
      float t[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const int I = Iv[c];
        const float xr = nxt(v[l], c);
        t[c] = (I <= IMM1) ? d.FP[I - 1] * (v[l][c] + xr) : 0.0f;
      }
      const float a = 0.5f * line_sum(t, IMM1, p.EXACT);
#pragma unroll
      for (int c = 0; c < NC; c++) v[l][c] = a;
    }
  }
  const unsigned ob = (unsigned)(J - 1) * (unsigned)IM + (unsigned)lane;
#pragma unroll
  for (int l = 0; l < L; l++) {
    const int K = kbase + l;
#pragma unroll
    for (int c = 0; c < NC; c++)
      if (K <= KM && Iv[c] <= IM) DST[ob + 64u * c + (unsigned)(K - 1) * plane] = v[l][c];
  }
}

// k_smooth_reg<L>w<NW>: NW wavefronts per J plane, L K lines each
// batched over the variables of e (grid z): q.DSC smoothed into q.D
#define SMOOTH_REG(L, NW, NC, SFX)                                                                           \
  __global__ void __launch_bounds__(64 * NW) k_smooth_reg##L##w##NW##SFX(Geo g, Dev d, EqSet e, Params p,    \
                                                                        StepScalars s, int j0) {             \
    if (gated(d)) return;                                                                                    \
    int z;                                                                                                   \
    const int J = j0 + (int)gridDim.x - 1 - eq_block(z); /* last planes (exit smoothing) first */             \
    const EqLane &q = eq_pick(e, z);                                                                         \
    __shared__ float lds[smooth_reg_lds<L, NW, NC>()];                                                       \
    smooth_reg<L, NW, NC>(g, d, q.DSC, q.D, p, s, q.ncall, J, lds);                                          \
  }
SMOOTH_REG(1, 8, 1, )
SMOOTH_REG(2, 8, 1, )
SMOOTH_REG(4, 8, 1, )
SMOOTH_REG(8, 8, 1, )
SMOOTH_REG(16, 8, 1, )
SMOOTH_REG(1, 16, 1, )
SMOOTH_REG(2, 16, 1, )
SMOOTH_REG(4, 16, 1, )
SMOOTH_REG(8, 16, 1, )
// two points per lane (IM <= 128)
SMOOTH_REG(1, 8, 2, c2)
SMOOTH_REG(2, 8, 2, c2)
SMOOTH_REG(4, 8, 2, c2)
SMOOTH_REG(8, 8, 2, c2)
SMOOTH_REG(16, 8, 2, c2)
SMOOTH_REG(1, 16, 2, c2)
SMOOTH_REG(2, 16, 2, c2)
SMOOTH_REG(4, 16, 2, c2)
SMOOTH_REG(8, 16, 2, c2)

#undef SMOOTH_REG

// ---------------------------------------------------------------------------
// LOSS, mf:8058-9066: mixing-length eddy viscosity and log-law wall friction
// turned into the body forces XFORCE, RFORCE, TFORCE, QSOURCE by a thin
// shear layer approximation.  YPLUSWALL >= -0.001 (no Shih wall functions),
// KM > 2.  k_loss_span forms the stresses on the K = const faces (one thread
// per (I,J) column, streaming up the column three times), k_loss_pitch those
// on the I = const faces and the blade surfaces (one wavefront per (J,K)
// line, the line held in LDS).  The entropy-generation and Reynolds-number
// diagnostics of the reference feed printed output only and are not formed.

// laminar viscosity and FTCOND while NSTEP = 1 or NSTEP < NSTEPS_MAX/4,
// mf:8082-8109 (one thread)
// Shih et al. wall function (NASA/TM-1999-209398), mf:15980-16047: the wall
// shear stress from the y+ of the previous call (*yplus, replaced by the
// new one, capped at 1000 as every caller does), with the streamwise
// pressure-gradient velocity when YPLUSWALL <= -10.  There |DPDS| < 1 is
// taken as 1 and stored back (the reference's DPDS is a by-reference
// argument); only the thread passing a non-null dpds_out stores.
__device__ float wallfun(float YPLUSWALL, float VISLAM, float PERP, float DPDS, float *dpds_out, float DENSITY,
                         float *yplus, float WREL_WALL, bool store) {
  const float YPLUS_OLD = *yplus;
  float SGN_DPDS = 0.0f, VPRES = 0.0f, VSTAR = 0.0f, FYPLUSP = 0.0f, FYPLUS = 0.0f, FRACP;
  if (!(YPLUSWALL > -10.0f)) {
    if (fabsf(DPDS) < 1.0f) {
      DPDS = 1.0f;
      if (store) *dpds_out = 1.0f;
    }
    SGN_DPDS = fabsf(DPDS) / DPDS;
    VPRES = pow_r(VISLAM / DENSITY / DENSITY * fabsf(DPDS), 0.33333f);
    VSTAR = YPLUS_OLD * VISLAM / (DENSITY * PERP);
    if (SGN_DPDS < 0.0f && VPRES > 0.33f * VSTAR) VPRES = 0.33f * VSTAR;
    if (SGN_DPDS > 0.0f && VPRES > 2.0f * VSTAR) VPRES = 2.0f * VSTAR;
    const float YP1 = VPRES * PERP * DENSITY / VISLAM, YP2 = YP1 * YP1, YP3 = YP2 * YP1, YP4 = YP2 * YP2;
    if (YP1 <= 4.0f) FYPLUSP = 0.5f * YP2 - 7.31E-3f * YP3;
    if (YP1 > 4.0f && YP1 <= 15.0f)
      FYPLUSP = -15.138f + 8.4688f * YP1 - 0.81976f * YP2 + 3.7292E-2f * YP3 - 6.3866E-4f * YP4;
    if (YP1 > 15.0f && YP1 <= 30.0f)
      FYPLUSP = 11.925f + 0.934f * YP1 - 2.7805E-2f * YP2 + 4.6262E-4f * YP3 - 3.1442E-6f * YP4;
    if (YP1 > 30.0f) FYPLUSP = 5.0f * log_r(YP1) + 8.0f;
  }
  const float YP1 = YPLUS_OLD, YP2 = YP1 * YP1, YP3 = YP2 * YP1, YP4 = YP2 * YP2;
  if (YP1 <= 5.0f) FYPLUS = YP1 + 1.0E-2f * YP2 - 2.9E-03f * YP3;
  if (YP1 > 5.0f && YP1 <= 30.0f) FYPLUS = -0.872f + 1.465f * YP1 - 7.02E-2f * YP2 + 1.66E-03f * YP3 - 1.495E-5f * YP4;
  if (YP1 > 30.0f && YP1 <= 140.0f)
    FYPLUS = 8.6f + 0.1864f * YP1 - 2.006E-3f * YP2 + 1.144E-5f * YP3 - 2.551E-8f * YP4;
  if (YP1 > 140.0f) FYPLUS = 2.439f * log_r(YP1) + 5.0f;
  if (YPLUSWALL > -10.0f) FRACP = 0.0f;
  else FRACP = SGN_DPDS * (VPRES / VSTAR) * FYPLUSP / FYPLUS;
  const float VPLUS = FYPLUS * (1.0f + FRACP);
  VSTAR = WREL_WALL / VPLUS;
  const float TWALL = DENSITY * VSTAR * VSTAR;
  if (store) *yplus = fminf(1000.0f, DENSITY * VSTAR * PERP / VISLAM);
  return TWALL;
}

__global__ void k_loss_vislam(Geo g, Dev d, Params p) {
  if (gated(d)) return;
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  float VISLAM = d.SC[SC_VISLAM];
  if (p.REYNO > 100.f) {
    int J1 = d.JLE[0], J2 = d.JTE[0];
    if (d.JLE[0] > g.JM) J1 = 1;
    if (d.JTE[0] > g.JM) J2 = g.JM;
    const float XCHORD = d.SMERID[JKX(J2, p.KMID)] - d.SMERID[JKX(J1, p.KMID)];
    const unsigned n = IX(p.IMID, J2, p.KMID);
    const float a = d.ROVX[n], b = d.ROVR[n], c = d.ROWT[n];
    const float ROW2 = sqrtf(a * a + b * b + c * c);
    VISLAM = XCHORD * ROW2 / p.REYNO;
  }
  if (p.REYNO > 0.0f && p.REYNO < 99.99f) VISLAM = p.REYNO / 100000.f;
  if (p.REYNO < 0.0f) VISLAM = -p.REYNO * 1.0e-5f;
  const float TCOND = p.CP * VISLAM / p.PRANDTL;
  d.SC[SC_VISLAM] = VISLAM;
  d.SC[SC_FTCOND] = TCOND / VISLAM;
}

// Skin-friction coefficient of a wall cell, mf:8287-8329 (log-law curve fit,
// roughness correction when ROUGH > 1e-7, at least laminar).  RO and WB are
// the density and speed the wall stress is formed with.
__device__ float loss_cf(float RE, float PERP, float ROUGH, float RO, float WB, float VISLAM) {
  const float RELOG = 1.0f / log_r(RE);
  float CF;
  if (ROUGH > 1.0e-7f) {
    if (ROUGH > PERP) ROUGH = PERP;
    float REK = RE * ROUGH / PERP;
    REK = REK - 80.0f;
    if (REK < 0.0f) REK = 0.0f;
    const float REKSQ = REK * REK;
    const float A1 = -.00178493f + .0000814923f * REK + .000000150445f * REKSQ;
    const float A2 = .029072f - .001584f * REK - .00000225194f * REKSQ;
    const float A3 = .270313f + .0091409f * REK + .00000451537f * REKSQ + .00000000464767f * REKSQ * REK;
    CF = A1 + A2 * RELOG + A3 * RELOG * RELOG;
  } else {
    CF = -0.00178493f + 0.029072f * RELOG + 0.270313f * RELOG * RELOG;
  }
  const float CFLAM = 2.0f / RE;
  CF = fmaxf(CF, CFLAM);
  if (RE < 125.f) CF = CFLAM;
  if (ROUGH > 1.0e-7f) {
    const float TW = 0.5f * CF * RO * WB * WB;
    const float VSTAR = sqrtf(TW / RO);
    const float PLUSK = ROUGH * VSTAR * RO / VISLAM;
    const float RELOG10 = log10f(PERP / ROUGH);
    const float FUNCN = 5.75f * RELOG10 + 8.5f;
    const float CF_FULL_ROUGH = 2.0f / (FUNCN * FUNCN);
    if (PLUSK > 45.f && CF > CF_FULL_ROUGH) CF = CF_FULL_ROUGH;
  }
  return CF;
}

// turbulent viscosity damping in the sublayer and buffer region, mf:8496-8501
__device__ __forceinline__ float loss_fyplus(float yp, float YPLAM, float YPTURB, float prev) {
  float f = prev;
  if (yp <= YPLAM) f = 0.0f;
  if (yp > YPLAM) {
    float XFAC = (yp - YPLAM) / (YPTURB - YPLAM);
    if (XFAC > 1.0f) XFAC = 1.0f;
    f = XFAC * XFAC * (3.0f - 2.0f * XFAC);
  }
  return f;
}

// face averages of one K level of an (I,J) column, mf:8220-8235
struct LossQ {
  float vx, vr, vt, ro, t, area, ravg, wabs;
};

__device__ __forceinline__ LossQ loss_q_span(const Geo &g, const Dev &d, int I, int J, int K, float WREL) {
  LossQ q;
  const unsigned a = IX(I, J, K), b = IX(I, J - 1, K), c = IX(I + 1, J, K), e = IX(I + 1, J - 1, K);
  const float ax = d.ASX[a], ar = d.ASR[a];
  q.area = sqrtf(ax * ax + ar * ar);
  q.vx = 0.25f * (d.VX[a] + d.VX[b] + d.VX[c] + d.VX[e]);
  q.vr = 0.25f * (d.VR[a] + d.VR[b] + d.VR[c] + d.VR[e]);
  q.vt = 0.25f * (d.VT[a] + d.VT[b] + d.VT[c] + d.VT[e]);
  q.ro = 0.25f * (d.RO[a] + d.RO[b] + d.RO[c] + d.RO[e]);
  q.t = 0.25f * (d.T_STATIC[a] + d.T_STATIC[b] + d.T_STATIC[c] + d.T_STATIC[e]);
  q.ravg = 0.5f * (d.R[JKX(J, K)] + d.R[JKX(J - 1, K)]);
  const float wt = q.vt - WREL * q.ravg;
  q.wabs = sqrtf(q.vx * q.vx + q.vr * q.vr + wt * wt);
  return q;
}

// speed relative to the end wall, mf:8231-8234
__device__ __forceinline__ float loss_wbound(const LossQ &q, float W) {
  const float wtb = q.vt - W * q.ravg;
  return sqrtf(q.vx * q.vx + q.vr * q.vr + wtb * wtb);
}

struct LossS {
  float x, r, t, q, w;
};

// Stresses on the K = const faces, forces relaxed by RF_VIS, mf:8201-8556.
__global__ void k_loss_span(Geo g, Dev d, Params p, float RF_VIS) {
  if (gated(d)) return;
  const float FMIXUP = d.SC[SC_FMIXUP];
  const int I = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int J = blockIdx.y * blockDim.y + threadIdx.y + 2;
  if (I > g.IMM1 || J > g.JM) return;
  const int KM = g.KM, KMM1 = g.KMM1, KMID = p.KMID;
  const int NRW = d.NROW[J - 1], JROW = J - d.JSTART[NRW - 1] + 1;
  const float WREL = d.WRAD[J - 1], WHUB = d.WHUB[J - 1], WTIP = d.WTIP[J - 1];
  const float RF_VIS1 = 1.f - RF_VIS;
  // end-wall friction, mf:8238-8418
  const LossQ q1 = loss_q_span(g, d, I, J, 1, WREL), q2 = loss_q_span(g, d, I, J, 2, WREL);
  const LossQ qm1 = loss_q_span(g, d, I, J, KMM1, WREL), qm = loss_q_span(g, d, I, J, KM, WREL);
  float VISLAM = d.SC[SC_VISLAM], FTCOND = d.SC[SC_FTCOND];
  if (p.REYNO < 0.0001f) {
    const LossQ qk = loss_q_span(g, d, I, J, KMID, WREL);
    VISLAM = (fabsf(p.REYNO) / 100000.0f) * pow_r(qk.t / 288.0f, 0.62f);
    const float TCOND = p.CP * VISLAM / p.PRANDTL;
    FTCOND = TCOND / VISLAM;
  }
  const float wb1 = loss_wbound(q1, WHUB), wb2 = loss_wbound(q2, 2 > KMID ? WTIP : WHUB);
  const float wbm1 = loss_wbound(qm1, KMM1 > KMID ? WTIP : WHUB), wbm = loss_wbound(qm, WTIP);
  float TWALLK1, TWALLKM;
  if (p.YPLUSWALL < -0.001f) {
    // Shih wall functions, mf:8269-8280 and 8348-8358
    const unsigned c1 = IX(I, J, 1), cm = IX(I, J, KMM1);
    TWALLK1 = wallfun(p.YPLUSWALL, VISLAM, d.VOL[c1] / q1.area, d.DPDS_CELL[c1], &d.DPDS_CELL[c1], q1.ro,
                      &d.YPLUS_K1[IJX(I, J)], wb1, true);
    TWALLKM = wallfun(p.YPLUSWALL, VISLAM, d.VOL[cm] / qm.area, d.DPDS_CELL[cm], &d.DPDS_CELL[cm], qm1.ro,
                      &d.YPLUS_KM[IJX(I, J)], wbm1, true);
  } else if (p.YPLUSWALL < 5.f) {
    const float PERP = d.VOL[IX(I, J, 1)] / q1.area;
    const float RE = PERP * q2.ro * wb2 / VISLAM;
    const float CF = loss_cf(RE, PERP, d.ROUGH_H[NRW - 1], q2.ro, wb2, VISLAM);
    TWALLK1 = 0.5f * CF * q2.ro * wb2 * wb2;
    const float PERPM = d.VOL[IX(I, J, KMM1)] / qm.area;
    const float REM = PERPM * qm1.ro * wbm1 / VISLAM;
    const float CFM = loss_cf(REM, PERPM, d.ROUGH_T[NRW - 1], qm1.ro, wbm1, VISLAM);
    TWALLKM = 0.5f * CFM * qm1.ro * wbm1 * wbm1;
  } else {
    TWALLK1 = p.CFWALL * q1.ro * wb1 * wb1;
    TWALLKM = p.CFWALL * qm.ro * wbm * wbm;
  }
  const float VSTAR1 = sqrtf(TWALLK1 / q1.ro), VSTARM = sqrtf(TWALLKM / qm.ro);
  LossS s1, sm;
  {
    float FMULT = TWALLK1 * q1.area / wb1;
    if (p.IBOUND == 1 || p.IBOUND > 2) FMULT = 0.0f;
    s1.x = FMULT * q1.vx;
    s1.r = FMULT * q1.vr;
    s1.t = FMULT * (q1.vt - WHUB * q1.ravg);
    FMULT = -TWALLKM * qm.area / wbm;
    if (p.IBOUND >= 2) FMULT = 0.0f;
    sm.x = FMULT * qm.vx;
    sm.r = FMULT * qm.vr;
    sm.t = FMULT * (qm.vt - WTIP * qm.ravg);
    s1.w = s1.t * WHUB * q1.ravg;
    s1.q = 0.0f;
    sm.w = sm.t * WTIP * qm.ravg;
    sm.q = 0.0f;
  }
  // y+ from the end walls (mf:8420-8435), the viscosities and the
  // transition test (mf:8437-8500)
  auto yplus = [&](int K) -> float {
    const float PERP = sqrtf(d.DWALLSQ[IX(I, J, K)]);
    return (K <= KMID) ? fminf(VSTAR1 * PERP * q1.ro / VISLAM, 1000.0f) : fminf(VSTARM * PERP * qm.ro / VISLAM, 1000.0f);
  };
  d.Y_PLUS[IX(I, J, 1)] = yplus(1);
  float RMAXH = 0.0f, RMAXT = 0.0f, RKMID = 0.0f;
  {
    LossQ qa = q1, qb = q2;
    float fy = 0.0f;
    for (int K = 2; K <= KMM1; K++) {
      const LossQ qc = loss_q_span(g, d, I, J, K + 1, WREL);
      const float yp = yplus(K);
      d.Y_PLUS[IX(I, J, K)] = yp;
      fy = loss_fyplus(yp, p.YPLAM, p.YPTURB, fy);
      const float DPERP = d.VOL[IX(I, J, K)] / qb.area;
      const float VLAM = d.LTAB[lt_span(LT_FKLAM, NRW, K, g.NROWS, g.IM, KM)] * VISLAM / DPERP;
      float VTURB = d.LTAB[lt_span(LT_FKTURB, NRW, K, g.NROWS, g.IM, KM)] * qb.ro * fabsf(qc.wabs - qa.wabs) * FMIXUP;
      VTURB = VTURB * fy;
      const float RATVIS = VTURB / VLAM;
      if (K <= KMID && !(RATVIS < RMAXH)) RMAXH = RATVIS;
      if (K == KMID) RKMID = RATVIS;
      if (K > KMID && !(RATVIS < RMAXT)) RMAXT = RATVIS;
      qa = qb;
      qb = qc;
    }
  }
  const bool keep_hub = (JROW >= d.JTRAN_K1[NRW - 1]) && (RMAXH > p.FTRANS);
  if (keep_hub && !(RKMID < RMAXT)) RMAXT = RKMID;  // the casing test re-reads VTURB(KMID)
  const bool keep_tip = (JROW >= d.JTRAN_KM[NRW - 1]) && (RMAXT > p.FTRANS);
  // stresses (mf:8503-8522) and forces (mf:8526-8535), streaming up in K
  LossS sk = s1;
  float ravgk = q1.ravg;
  LossQ qa = q1, qb = q2;
  float fy = 0.0f;
  for (int K = 1; K <= KMM1; K++) {
    LossS sn;
    float ravgn;
    if (K + 1 <= KMM1) {
      const int KN = K + 1;
      const LossQ qc = loss_q_span(g, d, I, J, KN + 1, WREL);
      const float yp = yplus(KN);
      fy = loss_fyplus(yp, p.YPLAM, p.YPTURB, fy);
      const float DPERP = d.VOL[IX(I, J, KN)] / qb.area;
      const float VLAM = d.LTAB[lt_span(LT_FKLAM, NRW, KN, g.NROWS, g.IM, KM)] * VISLAM / DPERP;
      float VTURB = d.LTAB[lt_span(LT_FKTURB, NRW, KN, g.NROWS, g.IM, KM)] * qb.ro * fabsf(qc.wabs - qa.wabs) * FMIXUP;
      VTURB = VTURB * fy;
      if ((KN <= KMID && !keep_hub) || (KN >= KMID && !keep_tip)) VTURB = 0.0f;
      const float FM = qb.area * (VLAM + VTURB);
      sn.x = FM * (qc.vx - qa.vx);
      sn.r = FM * (qc.vr - qa.vr);
      const float VISTOT = VISLAM * (1.0f + VTURB / VLAM);
      sn.t = FM * (qc.vt - qa.vt) - VISTOT * qb.area * qb.vt / qb.ravg;
      sn.q = FM * FTCOND * (qc.t - qa.t);
      sn.w = sn.x * qb.vx + sn.t * qb.vt + sn.r * qb.vr;
      ravgn = qb.ravg;
      qa = qb;
      qb = qc;
    } else {
      sn = sm;
      ravgn = qm.ravg;
    }
    const unsigned n = IX(I, J, K);
    d.XFORCE[n] = RF_VIS1 * d.XFORCE[n] + RF_VIS * (sk.x - sn.x);
    d.RFORCE[n] = RF_VIS1 * d.RFORCE[n] + RF_VIS * (sk.r - sn.r);
    d.TFORCE[n] = RF_VIS1 * d.TFORCE[n] + RF_VIS * (sk.t * ravgk - sn.t * ravgn);
    d.QSOURCE[n] = RF_VIS1 * d.QSOURCE[n] + RF_VIS * (sk.q - sn.q + sk.w - sn.w);
    sk = sn;
    ravgk = ravgn;
  }
}

// wavefront max of x over lanes (non-NaN values; the order does not matter)
__device__ __forceinline__ float wave_max(float x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const float y = __shfl_xor(x, off, 64);
    x = (y > x) ? y : x;
  }
  return x;
}

// Stresses on the blade surfaces and the I = const faces, added to the
// forces, mf:8562-9001.  One 64-thread workgroup per (J,K) line.
constexpr int kLossLine = 260;  // IM <= 257

__global__ void __launch_bounds__(64) k_loss_pitch(Geo g, Dev d, Params p, float RF_VIS) {
  if (gated(d)) return;
  const float FMIXUP = d.SC[SC_FMIXUP];
  __shared__ float AREA[kLossLine], VXA[kLossLine], VRA[kLossLine], VTA[kLossLine], WTA[kLossLine];
  __shared__ float ROA[kLossLine], TA[kLossLine], WABS[kLossLine];
  __shared__ float SX[kLossLine], SR[kLossLine], ST[kLossLine], SQ[kLossLine], SW[kLossLine];
  __shared__ float VLM[kLossLine], VTU[kLossLine], YP[kLossLine];
  const int K = blockIdx.x + 1, J = blockIdx.y + 2;
  const int IM = g.IM, IMM1 = g.IMM1, IMID = p.IMID, lane = threadIdx.x;
  const int NRW = d.NROW[J - 1], JROW = J - d.JSTART[NRW - 1] + 1;
  const int JLEDGE = d.JLE[NRW - 1], JTEDGE = d.JTE[NRW - 1];
  const float WREL = d.WRAD[J - 1], RAVGK = d.RAVG_CELL[JKX(J, K)], abt = d.ABT[JKX(J, K)];
  const int kts = d.KTIPS[NRW - 1];
  const bool tipgap = kts > 0 && K >= kts && K < d.KTIPE[NRW - 1];
  for (int I = lane + 1; I <= IM; I += 64) {
    const unsigned a = IX(I, J, K), b = IX(I, J - 1, K), c = IX(I, J - 1, K + 1), e = IX(I, J, K + 1);
    const float bx = d.ABX[a], br = d.ABR[a];
    AREA[I] = sqrtf(bx * bx + br * br + abt * abt);
    const float vx = 0.25f * (d.VX[a] + d.VX[b] + d.VX[c] + d.VX[e]);
    const float vr = 0.25f * (d.VR[a] + d.VR[b] + d.VR[c] + d.VR[e]);
    const float wt = 0.25f * (d.WT[a] + d.WT[b] + d.WT[c] + d.WT[e]);
    VXA[I] = vx;
    VRA[I] = vr;
    VTA[I] = 0.25f * (d.VT[a] + d.VT[b] + d.VT[c] + d.VT[e]);
    WTA[I] = wt;
    ROA[I] = 0.25f * (d.RO[a] + d.RO[b] + d.RO[c] + d.RO[e]);
    TA[I] = 0.25f * (d.T_STATIC[a] + d.T_STATIC[b] + d.T_STATIC[c] + d.T_STATIC[e]);
    WABS[I] = sqrtf(vx * vx + vr * vr + wt * wt);
  }
  __syncthreads();
  float VISLAM = d.SC[SC_VISLAM], FTCOND = d.SC[SC_FTCOND];
  if (p.REYNO < 0.0001f) {
    VISLAM = (fabsf(p.REYNO) / 100000.0f) * pow_r(TA[IMID] / 288.0f, 0.62f);
    const float TCOND = p.CP * VISLAM / p.PRANDTL;
    FTCOND = TCOND / VISLAM;
  }
  if (!(J <= JLEDGE || J > JTEDGE)) {
    // blade-surface friction and y+, mf:8633-8807
    float VSTAR1 = 0.0f, VSTARM = 0.0f;
    if (!tipgap) {
      float TWALLI1, TWALLIM;
      if (p.YPLUSWALL < -0.001f) {
        // Shih wall functions, mf:8621-8631 and 8696-8706; every lane
        // forms the stress, lane 0 keeps the wall state
        const unsigned c1 = IX(1, J, K), cm = IX(IMM1, J, K);
        const float dp1 = d.DPDS_CELL[c1], dpm = d.DPDS_CELL[cm];
        float yp1 = d.YPLUS_I1[JKX(J, K)], ypm = d.YPLUS_IM[JKX(J, K)];
        const bool keep = lane == 0;
        TWALLI1 = wallfun(p.YPLUSWALL, VISLAM, d.VOL[c1] / AREA[1], dp1, &d.DPDS_CELL[c1], ROA[1], &yp1, WABS[2], keep);
        TWALLIM = wallfun(p.YPLUSWALL, VISLAM, d.VOL[cm] / AREA[IM], dpm, &d.DPDS_CELL[cm], ROA[IMM1], &ypm,
                          WABS[IMM1], keep);
        if (keep) {
          d.YPLUS_I1[JKX(J, K)] = yp1;
          d.YPLUS_IM[JKX(J, K)] = ypm;
        }
      } else if (p.YPLUSWALL < 5.f) {
        const float PERP = d.VOL[IX(1, J, K)] / AREA[1];
        const float RE = ROA[2] * WABS[2] * PERP / VISLAM;
        const float CF = loss_cf(RE, PERP, d.ROUGH_L[NRW - 1], ROA[2], WABS[2], VISLAM);
        TWALLI1 = 0.5f * CF * ROA[2] * WABS[2] * WABS[2];
        const float PERPM = d.VOL[IX(IMM1, J, K)] / AREA[IM];
        const float REM = ROA[IMM1] * WABS[IMM1] * PERPM / VISLAM;
        const float CFM = loss_cf(REM, PERPM, d.ROUGH_U[NRW - 1], ROA[IMM1], WABS[IMM1], VISLAM);
        TWALLIM = 0.5f * CFM * ROA[IMM1] * WABS[IMM1] * WABS[IMM1];
      } else {
        TWALLI1 = p.CFWALL * ROA[1] * WABS[1] * WABS[1];
        TWALLIM = p.CFWALL * ROA[IM] * WABS[IM] * WABS[IM];
      }
      VSTAR1 = sqrtf(TWALLI1 / ROA[1]);
      VSTARM = sqrtf(TWALLIM / ROA[IM]);
      if (lane == 0) {
        float FM = TWALLI1 * AREA[1] / WABS[1];
        SX[1] = FM * VXA[1];
        SR[1] = FM * VRA[1];
        ST[1] = FM * WTA[1];
        FM = -TWALLIM * AREA[IM] / WABS[IM];
        SX[IM] = FM * VXA[IM];
        SR[IM] = FM * VRA[IM];
        ST[IM] = FM * WTA[IM];
        SQ[1] = 0.0f;
        SW[1] = ST[1] * WREL * RAVGK;
        SQ[IM] = 0.0f;
        SW[IM] = ST[IM] * WREL * RAVGK;
      }
    }
    // y+ including the blade surfaces (mf:8768-8783), kept in LDS
    for (int I = lane + 1; I <= IMM1; I += 64) {
      const unsigned n = IX(I, J, K);
      float yp = d.Y_PLUS[n];
      if (!tipgap) {
        const float YPLS = (I <= IMID) ? VSTAR1 * sqrtf(d.DWALLSQ[n]) * ROA[1] / VISLAM
                                       : VSTARM * sqrtf(d.DWALLSQ[n]) * ROA[IM] / VISLAM;
        yp = fminf(YPLS, yp);
        d.Y_PLUS[n] = yp;
      }
      YP[I] = yp;
    }
    __syncthreads();
    // viscosities and transition, mf:8813-8869
    float rlo = 0.0f, rhi = 0.0f, rmid = 0.0f;
    for (int I = lane + 2; I <= IMM1; I += 64) {
      const unsigned n = IX(I, J, K);
      const float fy = loss_fyplus(YP[I], p.YPLAM, p.YPTURB, 0.0f);
      const float DPERP = d.VOL[n] / AREA[I];
      const float VLAM = d.LTAB[lt_pitch(LT_FILAM, NRW, I, g.NROWS, IM)] * VISLAM / DPERP;
      float VTURB = d.LTAB[lt_pitch(LT_FITURB, NRW, I, g.NROWS, IM)] * ROA[I] * fabsf(WABS[I + 1] - WABS[I - 1]) * FMIXUP;
      VTURB = VTURB * fy;
      VLM[I] = VLAM;
      VTU[I] = VTURB;
      const float RATVIS = VTURB / VLAM;
      if (I <= IMID) rlo = (RATVIS > rlo) ? RATVIS : rlo;
      if (I == IMID) rmid = RATVIS;
      if (I > IMID) rhi = (RATVIS > rhi) ? RATVIS : rhi;
    }
    rlo = wave_max(rlo);
    rhi = wave_max(rhi);
    rmid = wave_max(rmid);  // only one lane holds I = IMID; the others hold 0
    const bool keep_lo = (JROW >= d.JTRAN_I1[NRW - 1]) && (rlo > p.FTRANS);
    if (keep_lo) rhi = (rmid > rhi) ? rmid : rhi;  // the upper test re-reads VTURB(IMID)
    const bool keep_hi = (JROW >= d.JTRAN_IM[NRW - 1]) && (rhi > p.FTRANS);
    // stresses, mf:8873-8883
    for (int I = lane + 2; I <= IMM1; I += 64) {
      float VTURB = VTU[I];
      if ((I <= IMID && !keep_lo) || (I >= IMID && !keep_hi)) VTURB = 0.0f;
      const float FM = (VLM[I] + VTURB) * AREA[I];
      SX[I] = FM * (VXA[I + 1] - VXA[I - 1]);
      SR[I] = FM * (VRA[I + 1] - VRA[I - 1]);
      ST[I] = FM * (WTA[I + 1] - WTA[I - 1]);
      SQ[I] = FM * FTCOND * (TA[I + 1] - TA[I - 1]);
      SW[I] = SX[I] * VXA[I] + SR[I] * VRA[I] + ST[I] * VTA[I];
    }
    __syncthreads();
    if (tipgap && lane == 0) {  // periodic stresses above the tip, mf:8889-8905
      SX[1] = 0.5f * (SX[2] + SX[IMM1]);
      SR[1] = 0.5f * (SR[2] + SR[IMM1]);
      ST[1] = 0.5f * (ST[2] + ST[IMM1]);
      SX[IM] = SX[1];
      ST[IM] = ST[1];
      SR[IM] = SR[1];
      SQ[1] = 0.5f * (SQ[2] + SQ[IMM1]);
      SW[1] = 0.5f * (SW[2] + SW[IMM1]);
      SQ[IM] = SQ[1];
      SW[IM] = SW[1];
    }
  } else {
    // upstream of the leading edge and in the wake, mf:8912-8977
    const bool wake = J > JTEDGE;
    float FACJ = 0.1f * (float)(J - JTEDGE);
    if (FACJ > 1.f) FACJ = 1.f;
    for (int I = lane + 1; I <= IM; I += 64) {
      const int ISUB = (I == IM) ? IMM1 : I;
      const int IM1 = (I == 1) ? IMM1 : I - 1, IP1 = (I == IM) ? 2 : I + 1;
      const float DPERP = d.VOL[IX(ISUB, J, K)] / AREA[I];
      const float VLAM = d.LTAB[lt_pitch(LT_FILAM, NRW, I, g.NROWS, IM)] * VISLAM / DPERP;
      float VTURB;
      if (wake) {
        const float FBLEND = (FACJ * d.LTAB[lt_pitch(LT_FIWAKE, NRW, I, g.NROWS, IM)] +
                              (1.f - FACJ) * d.LTAB[lt_pitch(LT_FITURB, NRW, I, g.NROWS, IM)]) * FMIXUP;
        VTURB = FBLEND * ROA[I] * fabsf(WABS[IP1] - WABS[IM1]);
      } else {
        VTURB = FMIXUP * d.LTAB[lt_pitch(LT_FIUP, NRW, I, g.NROWS, IM)] * ROA[I] * fabsf(WABS[IP1] - WABS[IM1]);
      }
      const float FM = AREA[I] * (VLAM + VTURB);
      const float sx = FM * (VXA[IP1] - VXA[IM1]), sr = FM * (VRA[IP1] - VRA[IM1]);
      const float st = FM * (WTA[IP1] - WTA[IM1]);
      SX[I] = sx;
      SR[I] = sr;
      ST[I] = st;
      SQ[I] = FM * FTCOND * (TA[IP1] - TA[IM1]);
      SW[I] = sx * VXA[I] + st * VTA[I] + sr * VRA[I];
    }
    __syncthreads();
    if (lane == 0) {  // periodic, mf:8982-8991
      SX[1] = 0.5f * (SX[IM] + SX[1]);
      SR[1] = 0.5f * (SR[IM] + SR[1]);
      ST[1] = 0.5f * (ST[IM] + ST[1]);
      SX[IM] = SX[1];
      SR[IM] = SR[1];
      ST[IM] = ST[1];
      SQ[1] = 0.5f * (SQ[IM] + SQ[1]);
      SW[1] = 0.5f * (SW[IM] + SW[1]);
      SQ[IM] = SQ[1];
      SW[IM] = SW[1];
    }
  }
  __syncthreads();
  // forces, mf:9005-9012
  for (int I = lane + 1; I <= IMM1; I += 64) {
    const unsigned n = IX(I, J, K);
    d.XFORCE[n] = d.XFORCE[n] + RF_VIS * (SX[I] - SX[I + 1]);
    d.RFORCE[n] = d.RFORCE[n] + RF_VIS * (SR[I] - SR[I + 1]);
    d.TFORCE[n] = d.TFORCE[n] + RF_VIS * (ST[I] - ST[I + 1]) * RAVGK;
    d.QSOURCE[n] = d.QSOURCE[n] + RF_VIS * (SQ[I] - SQ[I + 1] + SW[I] - SW[I + 1]);
  }
}

// ---------------------------------------------------------------------------
// Spalart-Allmaras model (ILOS >= 200): the per-step transport of the
// turbulent viscosity (mf:4260-4351, TSTEP NCALL=2 mf:5736-5777) and
// SPAL_LOSS (mf:13586-14827).  The turbulent viscosity is cell-centred.
// Restated without the Shih wall functions, FAC_PGRAD and FAC_VORT.

#define CELL_IJK \
  NODE_IJK       \
  if (I > g.IMM1 || K > g.KMM1 || J < 2) return;

// cell-average density, mf:4260-4270
__global__ void k_sa_cellavg(Geo g, Dev d) {
  if (gated(d)) return;
  CELL_IJK
  const float *RO = d.RO;
  d.ROAVG_CELL[IX(I, J, K)] = 0.125f * (RO[IX(I, J, K)] + RO[IX(I + 1, J, K)] + RO[IX(I + 1, J, K + 1)] +
                                        RO[IX(I, J, K + 1)] + RO[IX(I, J - 1, K)] + RO[IX(I + 1, J - 1, K)] +
                                        RO[IX(I + 1, J - 1, K + 1)] + RO[IX(I, J - 1, K + 1)]);
}

// doubled face fluxes and source of the turbulent viscosity, mf:4279-4330
// fx, ft, fr: the node's FLOWX, FLOWT, FLOWR; ft1: FLOWT(1,J,K)
__device__ __forceinline__ FaceFlux sa_faces(const Geo &g, const Dev &d, int I, int J, int K, float fx, float ft,
                                             float fr, float ft1) {
  const int IMM1 = g.IMM1;
  const float *TK = d.TRANS_KVIS;
  const unsigned n = IX(I, J, K);
  FaceFlux v;
  if (K == g.KM) {
    v.hr = J >= 2 && I <= IMM1;
    return v;
  }
  if (I <= IMM1) {
    int JP1 = J + 1;
    if (J == g.JM) JP1 = g.JM;
    if (J == 1) JP1 = 1;
    if (d.INDMIX[J - 1] == 1) JP1 = J;
    v.x = fx * (TK[n] + TK[IX(I, JP1, K)]);
    v.hx = true;
  }
  if (J < 2) return v;
  if (I <= IMM1) {
    v.s = -2.0f * d.T_SOURCE[n];
    v.r = (K == 1) ? 0.0f : fr * (TK[IX(I, J, K - 1)] + TK[n]);
    v.hs = v.hr = true;
  }
  if (I >= 2 && I <= IMM1) {
    v.t = ft * (TK[IX(I - 1, J, K)] + TK[n]);
    v.ht = true;
  }
  if (I == 1 || I == g.IM) {
    const unsigned n1 = IX(1, J, K);
    v.t = ft1 * (TK[IX(IMM1, J, K)] + TK[n1]);
    v.ht = true;
  }
  return v;
}
__device__ __forceinline__ void sa_fluxes_at(const Geo &g, const Dev &d, int I, int J, int K) {
  const size_t n = IX(I, J, K);
  store_faces(sa_faces(g, d, I, J, K, d.FLOWX[n], d.FLOWT[n], d.FLOWR[n], d.FLOWT[IX(1, J, K)]), d.XFLUX, d.RFLUX,
              d.TFLUX, d.SOURCE, n);
}
__global__ void k_sa_fluxes(Geo g, Dev d) {
  if (gated(d)) return;
  NODE_IJK
  sa_fluxes_at(g, d, I, J, K);
}

// D += STORE (damped as TSTEP's negative feedback leaves it) on the cells,
// mf:5716-5744
// The transport variable's TSTEP tail (the damped change added, the pitch
// average carried over each mixing plane, SMOOTH_RESID's three directional
// passes, the line conditions) in two line kernels: one wavefront per (J,K)
// line, four lines per workgroup along K, each pass's statements in the
// reference's order.
// k_sa_tail1: the new value D' = D + (damped) STORE (mf:5736-5746) of the
// lines J-1, J, J+1, with the pitch average carried over each mixing plane
// (mf:5748-5762: D' of lines JMIX+1 and JMIX+2 is FMULT times the average of
// line JMIX), then the streamwise and pitchwise smoothing passes of
// SMOOTH_RESID with FAC_SFVIS*SFT (mf:8700), into TMP (cells only).
template <bool BS>
__device__ __forceinline__ void sa_new_line(const Geo &g, const Dev &d, Hist<BS> STORE, const Params &p,
                                            const float *D, int damp, float DAMP, int J, int K, int lane, int nc,
                                            float (&v)[4]) {
  int jx = 0;
  const int role = g.NROWS > 1 ? mp_role(d, J, jx) : 0;
  const int JS = (role >= 2) ? jx : J;  // the line whose D' this line takes
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int I = 64 * c + lane + 1;
    v[c] = 0.0f;
    if (c >= nc || I > g.IMM1) continue;
    const unsigned n = IX(I, JS, K);
    const float x = D[n];
    if (JS < 2) {
      v[c] = x;
    } else {
      const float st = STORE.ld(n);
      v[c] = x + (damp ? damped(g, d, d.AVG, st, JS, DAMP) : st);
    }
  }
  if (role >= 2) {
    const float FMULT = d.TURBVIS_DAMP[d.NROW[jx - 1] - 1];
    float t[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int I = 64 * c + lane + 1;
      t[c] = (c < nc && I <= g.IMM1) ? d.FP[I - 1] * v[c] : 0.0f;
    }
    const float AVG = line_sum(t, g.IMM1, p.EXACT);
#pragma unroll
    for (int c = 0; c < 4; c++) v[c] = FMULT * AVG;
  }
}
template <bool BS>
__device__ __forceinline__ void sa_tail1(const Geo &g, const Dev &d, const Params &p, const float *D, int damp,
                                         Hist<BS> S) {
  __shared__ float sh[4][260];
  const int lane = threadIdx.x & 63, w = wave_id();
  const int K = blockIdx.x * 4 + w + 1, J = blockIdx.y + 1;
  if (K > g.KMM1) return;
  const int IMM1 = g.IMM1, IMM2 = g.IM - 2, nc = (IMM1 + 63) / 64;
  const float DAMP = d.SC[SC_DAMP];
  const float SMFAC = p.FAC_SFVIS * d.SC[SC_SFT];
  const float SFAC1 = 1.0f - SMFAC, SFACH = 0.5f * SMFAC;
  const int NR = d.NROW[J - 1];
  const int J1P1 = d.JSTART[NR - 1] + 1, J2 = d.JMIX[NR - 1];
  float s0[4], u[4];
  sa_new_line(g, d, S, p, D, damp, DAMP, J, K, lane, nc, s0);
  if (J < J1P1) {
#pragma unroll
    for (int c = 0; c < 4; c++) u[c] = s0[c];
  } else {
    // streamwise: J1+1 takes J1+2, JMIX takes JMIX-1, others J-1 and J+1
    float a[4], b[4];
    if (J == J1P1) {
      sa_new_line(g, d, S, p, D, damp, DAMP, J + 1, K, lane, nc, a);
#pragma unroll
      for (int c = 0; c < 4; c++) u[c] = s0[c] + SFACH * (a[c] - s0[c]);
    } else if (J == J2) {
      sa_new_line(g, d, S, p, D, damp, DAMP, J - 1, K, lane, nc, a);
#pragma unroll
      for (int c = 0; c < 4; c++) u[c] = s0[c] + SFACH * (a[c] - s0[c]);
    } else {
      sa_new_line(g, d, S, p, D, damp, DAMP, J + 1, K, lane, nc, a);
      sa_new_line(g, d, S, p, D, damp, DAMP, J - 1, K, lane, nc, b);
#pragma unroll
      for (int c = 0; c < 4; c++) u[c] = SFAC1 * s0[c] + SFACH * (a[c] + b[c]);
    }
  }
  // pitchwise, from the line in LDS
  float *L = sh[w];
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int I = 64 * c + lane + 1;
    if (c < nc && I <= IMM1) L[I] = u[c];
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int I = 64 * c + lane + 1;
    if (c >= nc || I > IMM1) continue;
    float v;
    if (J < J1P1) v = L[I];
    else if (I == 1) v = L[I] + SFACH * (L[2] - L[I]);
    else if (I == IMM1) v = L[I] + SFACH * (L[IMM2] - L[I]);
    else v = SFAC1 * L[I] + SFACH * (L[I + 1] + L[I - 1]);
    d.TMP[IX(I, J, K)] = v;
  }
}
__global__ void __launch_bounds__(256) k_sa_tail1(Geo g, Dev d, Params p, const float *D, int damp) {
  if (gated(d)) return;
  sa_tail1<false>(g, d, p, D, damp, Hist<false>{d.STORE});
}
// the mixed form's fused chain: STORE from its BF16 copy
__global__ void __launch_bounds__(256) k_sa_tail1_bfs(Geo g, Dev d, Params p, const float *D, int damp,
                                                      unsigned short *S16) {
  if (gated(d)) return;
  sa_tail1<true>(g, d, p, D, damp, Hist<true>{S16});
}

// k_sa_tail2: the spanwise smoothing pass from TMP, then the periodic,
// tip-gap and mixing-plane conditions on the line with ILAST = IM-1
// (mf:5858-5926), into D; lines K = KM (no cells) take only the conditions.
__global__ void __launch_bounds__(256) k_sa_tail2(Geo g, Dev d, Params p, float *D) {
  if (gated(d)) return;
  __shared__ float sh[4][260];
  const int lane = threadIdx.x & 63, w = wave_id();
  const int K = blockIdx.x * 4 + w + 1, J = blockIdx.y + 1;
  if (K > g.KM) return;
  const int IM = g.IM, IMM1 = g.IMM1, KM = g.KM, KMM1 = g.KMM1, KMM2 = KM - 2;
  const float SMFAC = p.FAC_SFVIS * d.SC[SC_SFT];
  const float SFAC1 = 1.0f - SMFAC, SFACH = 0.5f * SMFAC;
  const int NR = d.NROW[J - 1];
  const int J1P1 = d.JSTART[NR - 1] + 1;
  float *L = sh[w];
  for (int I = lane + 1; I <= IM; I += 64) {
    const unsigned n = IX(I, J, K);
    float v;
    if (K > KMM1 || I > IMM1) {
      v = D[n];
    } else {
      const float *src = d.TMP;
      if (J < J1P1) v = src[n];
      else if (K == 1) v = src[n] + SFACH * (src[IX(I, J, 2)] - src[n]);
      else if (K == KMM1) v = src[n] + SFACH * (src[IX(I, J, KMM2)] - src[n]);
      else v = SFAC1 * src[n] + SFACH * (src[IX(I, J, K + 1)] + src[IX(I, J, K - 1)]);
    }
    L[I] = v;
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  const bool per = !((d.IND[J - 1] == 1) || (d.INDLE[J - 1] == 1));
  bool tip = false;
  if (d.KTIPS[NR - 1] > 0) {
    int K1 = d.KTIPS[NR - 1], K2 = d.KTIPE[NR - 1];
    if (K1 == 1) K2 = K2 - 1;
    if (K2 == KM) K1 = K1 + 1;
    tip = K >= K1 && K <= K2;
  }
  const bool mp = p.IFMIX != 0 && ((J == d.JMIX[NR - 1] && NR < g.NROWS) || (NR > 1 && J == d.JSTART[NR - 1]));
  if (lane == 0) {
    if (per) {
      L[1] = 0.5f * (L[1] + L[IMM1]);
      L[IMM1] = L[1];
    }
    if (tip) {
      L[1] = 0.5f * (L[1] + L[IMM1]);
      L[IMM1] = L[1];
    }
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  if (mp) {
    float v[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int I = 64 * c + lane + 1;
      v[c] = (I <= IMM1) ? d.FP[I - 1] * (L[I] + L[I + 1]) : 0.0f;
    }
    const float AVG = 0.5f * line_sum(v, IMM1, p.EXACT);
    for (int I = lane + 1; I <= IM; I += 64) D[IX(I, J, K)] = AVG;
  } else {
    for (int I = lane + 1; I <= IMM1; I += 64) D[IX(I, J, K)] = L[I];
  }
}

// lower limit and kinematic value of the turbulent viscosity, mf:4339-4348
__device__ __forceinline__ void sa_kvis_at(const Geo &g, const Dev &d, const Params &p, unsigned n, float roavg) {
  const float TRANSVISMIN = 0.1f * d.VISC_LAM[IX(p.IMID, 2, p.KMID)];
  float v = d.TRANS_DYN_VIS[n];
  if (v < TRANSVISMIN) v = TRANSVISMIN;
  d.TRANS_DYN_VIS[n] = v;
  d.TRANS_KVIS[n] = v / roavg;
}
__global__ void k_sa_kvis(Geo g, Dev d, Params p) {
  if (gated(d)) return;
  CELL_IJK
  const unsigned n = IX(I, J, K);
  sa_kvis_at(g, d, p, n, d.ROAVG_CELL[n]);
}
// the cell-average density (k_sa_cellavg) and then the above, per cell
__global__ void k_sa_cellavg_kvis(Geo g, Dev d, Params p) {
  if (gated(d)) return;
  CELL_IJK
  const float *RO = d.RO;
  const unsigned n = IX(I, J, K);
  const float roavg = 0.125f * (RO[IX(I, J, K)] + RO[IX(I + 1, J, K)] + RO[IX(I + 1, J, K + 1)] + RO[IX(I, J, K + 1)] +
                                RO[IX(I, J - 1, K)] + RO[IX(I + 1, J - 1, K)] + RO[IX(I + 1, J - 1, K + 1)] +
                                RO[IX(I, J - 1, K + 1)]);
  d.ROAVG_CELL[n] = roavg;
  sa_kvis_at(g, d, p, n, roavg);
}

// Gauss-theorem gradient of a corner-stored field over cell (I,J,K),
// mf:13468-13510
__device__ __forceinline__ void sa_gradvel(const Geo &g, const Dev &d, const float *V, int I, int J, int K,
                                           float &DVDX, float &DVDR, float &DVDT) {
  const int IP1 = I + 1, JM1 = J - 1, KP1 = K + 1;
  const float RVOL = 0.25f / d.VOL[IX(I, J, K)];
  const float VAVGJ1 = V[IX(I, JM1, K)] + V[IX(IP1, JM1, K)] + V[IX(IP1, JM1, KP1)] + V[IX(I, JM1, KP1)];
  const float VAVGJ2 = V[IX(I, J, K)] + V[IX(IP1, J, K)] + V[IX(IP1, J, KP1)] + V[IX(I, J, KP1)];
  const float VAVGI1 = V[IX(I, J, K)] + V[IX(I, JM1, K)] + V[IX(I, JM1, KP1)] + V[IX(I, J, KP1)];
  const float VAVGI2 = V[IX(IP1, J, K)] + V[IX(IP1, JM1, K)] + V[IX(IP1, JM1, KP1)] + V[IX(IP1, J, KP1)];
  const float VAVGK1 = V[IX(I, J, K)] + V[IX(IP1, J, K)] + V[IX(IP1, JM1, K)] + V[IX(I, JM1, K)];
  const float VAVGK2 = V[IX(I, J, KP1)] + V[IX(IP1, J, KP1)] + V[IX(IP1, JM1, KP1)] + V[IX(I, JM1, KP1)];
  const float VAVGCELL = 0.125f * (VAVGJ1 + VAVGJ2);
  DVDX = -(VAVGJ2 * d.AQX[IX(I, J, K)] - VAVGJ1 * d.AQX[IX(I, JM1, K)] + VAVGI1 * d.ABX[IX(I, J, K)] -
           VAVGI2 * d.ABX[IX(IP1, J, K)] + VAVGK1 * d.ASX[IX(I, J, K)] - VAVGK2 * d.ASX[IX(I, J, KP1)]) * RVOL;
  DVDR = -(VAVGJ2 * d.AQR[IX(I, J, K)] - VAVGJ1 * d.AQR[IX(I, JM1, K)] + VAVGI1 * d.ABR[IX(I, J, K)] -
           VAVGI2 * d.ABR[IX(IP1, J, K)] + VAVGK1 * d.ASR[IX(I, J, K)] - VAVGK2 * d.ASR[IX(I, J, KP1)]) * RVOL -
         VAVGCELL / d.RAVG_CELL[JKX(J, K)];
  DVDT = -(VAVGI1 - VAVGI2) * d.ABT[JKX(J, K)] * RVOL;
}

// Gradient of a cell-centred field, mf:13516-13580; cells beyond J = JM
// read as zero (the reference's unused, zero J = JM+1 layer).
__device__ __forceinline__ void sa_gradcell(const Geo &g, const Dev &d, const float *V, int I, int J, int K,
                                            int p_IBOUND, float &DVDX, float &DVDR, float &DVDT) {
  const int IMM1 = g.IMM1, KMM1 = g.KMM1;
  int IP1 = I + 1, JP1 = J + 1, KP1 = K + 1, IM1 = I - 1, JM1 = J - 1, KM1 = K - 1;
  if (I == 1) IM1 = IMM1;
  if (I == IMM1) IP1 = 1;
  if (K == 1) KM1 = 1;
  if (K == KMM1) KP1 = KMM1;
  const float RVOL = 0.5f / d.VOL[IX(I, J, K)];
  const float v0 = V[IX(I, J, K)];
  const float vjp = (JP1 <= g.JM) ? V[IX(I, JP1, K)] : 0.0f;
  float VAVGJ1 = v0 + V[IX(I, JM1, K)];
  float VAVGJ2 = v0 + vjp;
  float VAVGI1 = v0 + V[IX(IM1, J, K)];
  float VAVGI2 = v0 + V[IX(IP1, J, K)];
  float VAVGK1 = v0 + V[IX(I, J, KM1)];
  float VAVGK2 = v0 + V[IX(I, J, KP1)];
  if (d.IND[J - 1] == 1) {
    if (I == 1) VAVGI1 = 0.0f;
    if (I == IMM1) VAVGI2 = 0.0f;
  }
  if (K == 1 && (p_IBOUND == 0 || p_IBOUND == 2)) VAVGK1 = 0.0f;
  if (K == KMM1 && p_IBOUND <= 1) VAVGK2 = 0.0f;
  DVDX = -(VAVGJ2 * d.AQX[IX(I, J, K)] - VAVGJ1 * d.AQX[IX(I, JM1, K)] + VAVGI1 * d.ABX[IX(I, J, K)] -
           VAVGI2 * d.ABX[IX(IP1, J, K)] + VAVGK1 * d.ASX[IX(I, J, K)] - VAVGK2 * d.ASX[IX(I, J, KP1)]) * RVOL;
  DVDR = -(VAVGJ2 * d.AQR[IX(I, J, K)] - VAVGJ1 * d.AQR[IX(I, JM1, K)] + VAVGI1 * d.ABR[IX(I, J, K)] -
           VAVGI2 * d.ABR[IX(IP1, J, K)] + VAVGK1 * d.ASR[IX(I, J, K)] - VAVGK2 * d.ASR[IX(I, J, KP1)]) * RVOL -
         v0 / d.RAVG_CELL[JKX(J, K)];
  DVDT = -(VAVGI1 - VAVGI2) * d.ABT[JKX(J, K)] * RVOL;
}

// SET_PWALLGRAD, mf:16052-16114.  The streamwise pressure gradient of each
// cell (PEFF's Gauss gradient along the 8-node velocity sum), limited.
__global__ void k_pwallgrad_cell(Geo g, Dev d, Params p) {
  if (gated(d)) return;
  CELL_IJK
  const unsigned n = IX(I, J, K);
  // the limit: the mean pressure change per row over a tenth of the chord,
  // as the reference writes it (mf:16065-16067)
  float DPDS_LIM = 20.0f * (d.PO1[p.KMID - 1] - 0.5f * (p.PDOWN_HUB + p.PDOWN_TIP) / (float)g.NROWS /
                                                    (d.CHORD[0] + d.CHORD[g.NROWS - 1]));
  DPDS_LIM = fabsf(DPDS_LIM);
  float DPDX, DPDR, DPDT;
  sa_gradvel(g, d, d.PEFF, I, J, K, DPDX, DPDR, DPDT);
  const float *VX = d.VX, *VR = d.VR, *WT = d.WT;
  const float VXCELL = VX[IX(I, J, K)] + VX[IX(I, J - 1, K)] + VX[IX(I + 1, J, K)] + VX[IX(I + 1, J - 1, K)] +
                       VX[IX(I, J, K + 1)] + VX[IX(I, J - 1, K + 1)] + VX[IX(I + 1, J, K + 1)] + VX[IX(I + 1, J - 1, K + 1)];
  const float VRCELL = VR[IX(I, J, K)] + VR[IX(I, J - 1, K)] + VR[IX(I + 1, J, K)] + VR[IX(I + 1, J - 1, K)] +
                       VR[IX(I, J, K + 1)] + VR[IX(I, J - 1, K + 1)] + VR[IX(I + 1, J, K + 1)] + VR[IX(I + 1, J - 1, K + 1)];
  const float WTCELL = WT[IX(I, J, K)] + WT[IX(I, J - 1, K)] + WT[IX(I + 1, J, K)] + WT[IX(I + 1, J - 1, K)] +
                       WT[IX(I, J, K + 1)] + WT[IX(I, J - 1, K + 1)] + WT[IX(I + 1, J, K + 1)] + WT[IX(I + 1, J - 1, K + 1)];
  const float WREL_CELL = sqrtf(VXCELL * VXCELL + VRCELL * VRCELL + WTCELL * WTCELL);
  float DPDS = (DPDX * VXCELL + DPDR * VRCELL + DPDT * WTCELL) / WREL_CELL;
  if (fabsf(DPDS) > DPDS_LIM) DPDS = DPDS_LIM * DPDS / fabsf(DPDS);
  d.DPDS_CELL[n] = DPDS;
}

// The two smoothing passes along J within each row, with the row's end
// planes extrapolated before and after, mf:16089-16110; one thread per
// (I,K) line walks the rows in order (a row's end treatment writes the
// next row's second plane, which that row's smoothing reads).  Planes
// beyond JM read as zero and are not stored (the reference's unused, zero
// array padding).
__global__ void k_pwallgrad_smooth(Geo g, Dev d) {
  if (gated(d)) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.IMM1 * g.KMM1) return;
  const int I = t % g.IMM1 + 1, K = t / g.IMM1 + 1, JM = g.JM;
  float *D = d.DPDS_CELL;
  auto dp = [&](int J) -> float { return J <= JM ? D[IX(I, J, K)] : 0.0f; };
  auto set = [&](int J, float v) {
    if (J <= JM) D[IX(I, J, K)] = v;
  };
  for (int NR = 1; NR <= g.NROWS; NR++) {
    const int J1 = d.JSTART[NR - 1], J2 = d.JMIX[NR - 1];
    set(J2 + 1, 0.0f);
    set(J2 + 2, 2 * dp(J2 + 3) - dp(J2 + 4));
    set(J2, 2 * dp(J2 - 1) - dp(J2 - 2));
    // TEMPP(J) = 0.5 D(J) + 0.25 (D(J-1) + D(J+1)) for J1+2..J2-1, then
    // D(J) = 0.5 TEMPP(J) + 0.25 (TEMPP(J-1) + TEMPP(J+1)) for J1+3..J2-2,
    // streamed with a three-plane window of each.  Every D read is of a
    // plane not yet written, so the planes are fetched kCh at a time ahead
    // of the chunk's stores (one memory latency per chunk, not per plane;
    // planes past J2 are fetched but not used).
    constexpr int kCh = 16;
    float dm1 = dp(J1 + 1), d0 = dp(J1 + 2), dn = dp(J1 + 3), nx = dp(J1 + 4);
    float tm1 = 0.0f, t0 = 0.5f * d0 + 0.25f * (dm1 + dn);  // TEMPP(J1+2)
    for (int Jc = J1 + 3; Jc <= J2 - 2; Jc += kCh) {
      float w[kCh];  // D(J+2) of the chunk's planes J
#pragma unroll
      for (int q = 0; q < kCh; q++) w[q] = dp(Jc + q + 2);
#pragma unroll
      for (int q = 0; q < kCh; q++) {
        const int J = Jc + q;
        if (J <= J2 - 2) {
          dm1 = d0;
          d0 = dn;
          dn = nx;  // D(J+1)
          tm1 = t0;
          t0 = 0.5f * d0 + 0.25f * (dm1 + dn);            // TEMPP(J)
          const float tn = 0.5f * dn + 0.25f * (d0 + w[q]);  // TEMPP(J+1)
          D[IX(I, J, K)] = 0.5f * t0 + 0.25f * (tm1 + tn);
          nx = w[q];
        }
      }
    }
    set(J2 + 1, 0.0f);
    set(J2 + 2, 2 * dp(J2 + 3) - dp(J2 + 4));
    set(J2, 2 * dp(J2 - 1) - dp(J2 - 2));
  }
}

// velocity and temperature gradients, strain rates and vorticity of each
// cell, mf:13666-13718 (the vorticity before the mixing-plane substitution);
// NEW_LOSS forms the same quantities (mf:12586-12632) but keeps the
// vorticity of a row's first J, which SPAL_LOSS zeroes
__global__ void k_sa_grad(Geo g, Dev d, Params p) {
  if (gated(d)) return;
  CELL_IJK
  const unsigned n = IX(I, J, K);
  const int IP1 = I + 1, JM1 = J - 1, KP1 = K + 1;
  const float *WT = d.WT;
  const float WTAVG = 0.125f * (WT[IX(I, J, K)] + WT[IX(IP1, J, K)] + WT[IX(IP1, J, KP1)] + WT[IX(I, J, KP1)] +
                                WT[IX(I, JM1, K)] + WT[IX(IP1, JM1, K)] + WT[IX(IP1, JM1, KP1)] + WT[IX(I, JM1, KP1)]);
  const float RAVG = d.RAVG_CELL[JKX(J, K)];
  float DVXDX, DVXDR, DVXDT, DVRDX, DVRDR, DVRDT, DWTDX, DWTDR, DWTDT, DTDX, DTDR, DTDT;
  sa_gradvel(g, d, d.VX, I, J, K, DVXDX, DVXDR, DVXDT);
  sa_gradvel(g, d, d.VR, I, J, K, DVRDX, DVRDR, DVRDT);
  sa_gradvel(g, d, d.WT, I, J, K, DWTDX, DWTDR, DWTDT);
  sa_gradvel(g, d, d.T_STATIC, I, J, K, DTDX, DTDR, DTDT);
  d.SXR[n] = DVXDR + DVRDX;
  d.SXX[n] = DVXDX;
  d.SXT[n] = DWTDX + DVXDT;
  d.SRR[n] = DVRDR;
  d.SRT[n] = DWTDR + DVRDT - WTAVG / RAVG;
  d.STT[n] = DWTDT;
  d.GQX[n] = DTDX;
  d.GQR[n] = DTDR;
  d.GQT[n] = DTDT;
  const float VORTX = DVRDT - DWTDR - WTAVG / RAVG;
  const float VORTR = DWTDX - DVXDT;
  const float VORTT = DVXDR - DVRDX;
  float VORT = sqrtf(VORTX * VORTX + VORTR * VORTR + VORTT * VORTT);
  const int NRW = d.NROW[J - 1];
  if (p.ILOS >= 200 && J == d.JSTART[NRW - 1]) VORT = 0.0f;
  if (VORT > p.VORT_MAX) VORT = p.VORT_MAX;
  d.ABS_VORT[n] = VORT;
}

// turbulent viscosity of each cell, mf:13719-13888 (mixing-plane cells
// take the vorticity of the cell upstream; no two such J are adjacent)
// first: the call is the first step's (NSTEP = 1)
__global__ void k_sa_visc(Geo g, Dev d, Params p, int first) {
  if (gated(d)) return;
  const float FMIXUP = d.SC[SC_FMIXUP];
  const int NSTEP = first ? 1 : 2;
  CELL_IJK
  const unsigned n = IX(I, J, K);
  float VORT = d.ABS_VORT[n];
  if (d.INDMIX[J - 1] == 1) {
    VORT = d.ABS_VORT[IX(I, J - 1, K)];
    d.ABS_VORT[n] = VORT;
  }
  const float ROAVG = d.ROAVG_CELL[n];
  const float TURBVIS_WALL = ROAVG * d.XLENGTH[n] * VORT * FMIXUP;
  float VISLAM = d.SC[SC_VISLAM];
  if (p.REYNO < 0.0001f) {
    const int IP1 = I + 1, JM1 = J - 1, KP1 = K + 1;
    const float *T = d.T_STATIC;
    const float TSAVG = 0.125f * (T[IX(I, J, K)] + T[IX(IP1, J, K)] + T[IX(IP1, J, KP1)] + T[IX(I, J, KP1)] +
                                  T[IX(I, JM1, K)] + T[IX(IP1, JM1, K)] + T[IX(IP1, JM1, KP1)] + T[IX(I, JM1, KP1)]);
    VISLAM = (fabsf(p.REYNO) / 100000.0f) * pow_r(TSAVG / 288.0f, 0.62f);
  }
  float VISTOT = VISLAM * (1 + d.FST_RAT[n]) + TURBVIS_WALL;
  float VISLIM = p.TURBVIS_LIM * VISLAM;
  if (VISTOT > VISLIM) VISTOT = VISLIM;
  const float VISKLAM = VISLAM / ROAVG;
  d.DYN_VIS_ML[n] = (VISTOT - VISLAM);
  if (p.IF_RESTART == 0 && NSTEP == 1) {
    const float FST = d.FSTURB[0];
    const float TRANSVISIN = (FST < 1.0f) ? 4.35f * VISLAM * pow_r(FST, 0.25f) : VISLAM * (3.5f + FST * 0.85f);
    d.TRANS_DYN_VIS[n] = TRANSVISIN;
  }
  if (NSTEP == 1) d.TRANS_KVIS[n] = d.TRANS_DYN_VIS[n] / ROAVG;
  float TRANSKVIS = d.TRANS_KVIS[n];
  float XI = TRANSKVIS / VISKLAM;
  if (XI < 1.1f) XI = 1.1f;
  if (XI > p.TURBVIS_LIM) XI = p.TURBVIS_LIM;
  const float XI3 = XI * XI * XI;
  const float FV1 = XI3 / (XI3 + 357.9f);
  const float TRANSVISLIM = VISLIM / ROAVG / FV1;
  if (TRANSKVIS > TRANSVISLIM) TRANSKVIS = TRANSVISLIM;
  const float VISTRU = TRANSKVIS * FV1 * ROAVG;
  VISTOT = VISTRU + VISLAM;
  VISLIM = p.TURBVIS_LIM * VISLAM;
  if (VISTOT > VISLIM) VISTOT = VISLIM;
  d.VISC_RAT[n] = VISTOT / VISLAM;
  d.VISC_LAM[n] = VISLAM;
  d.VISC_TURB[n] = VISTOT - VISLAM;
}

// NEW_LOSS's viscosity of each cell, mf:12636-12672: the mixing length times
// the vorticity, damped through the buffer layer by the cell's y+ from the
// previous call, plus the laminar and free-stream-turbulence parts, limited
// to TURBVIS_LIM times laminar
__global__ void k_nl_visc(Geo g, Dev d, Params p) {
  if (gated(d)) return;
  const float FMIXUP = d.SC[SC_FMIXUP];
  CELL_IJK
  const unsigned n = IX(I, J, K);
  float TURBVIS_WALL = d.ROAVG_CELL[n] * d.XLENGTH[n] * d.ABS_VORT[n] * FMIXUP;
  const float yp = d.Y_PLUS[n];
  float FYPLUS = 0.0f;
  if (yp > p.YPLAM) {
    float XFAC = (yp - p.YPLAM) / (p.YPTURB - p.YPLAM);
    if (XFAC > 1.0f) XFAC = 1.0f;
    FYPLUS = XFAC * XFAC * (3.0f - 2.0f * XFAC);
  }
  TURBVIS_WALL = TURBVIS_WALL * FYPLUS;
  float VISLAM = d.SC[SC_VISLAM];
  if (p.REYNO < 0.0001f) {
    const int IP1 = I + 1, JM1 = J - 1, KP1 = K + 1;
    const float *T = d.T_STATIC;
    const float TSAVG = 0.125f * (T[IX(I, J, K)] + T[IX(IP1, J, K)] + T[IX(IP1, J, KP1)] + T[IX(I, J, KP1)] +
                                  T[IX(I, JM1, K)] + T[IX(IP1, JM1, K)] + T[IX(IP1, JM1, KP1)] + T[IX(I, JM1, KP1)]);
    VISLAM = (fabsf(p.REYNO) / 100000.0f) * pow_r(TSAVG / 288.0f, 0.62f);
  }
  float VISTOT = VISLAM * (1 + d.FST_RAT[n]) + TURBVIS_WALL;
  const float VISLIM = p.TURBVIS_LIM * VISLAM;
  if (VISTOT > VISLIM) VISTOT = VISLIM;
  d.VISC_RAT[n] = VISTOT / VISLAM;
  d.VISC_LAM[n] = VISLAM;
}

// NEW_LOSS's transition test, mf:12677-12772: a surface's cells within half
// the span (hub, casing) or half the pitch (I = 1 and I = IM surfaces) are
// laminar (viscosity ratio 1) upstream of its transition point and wherever
// the largest ratio there has not passed FTRANS.  One workgroup per J: the
// hub and casing tests of every (I) column come first, then the blade tests
// of every (K) line, which read what the column tests set.  The casing range
// starts at KMID and the I = IM range at IMID, so each re-reads the cell
// the test before it may have reset.
__global__ void __launch_bounds__(64) k_nl_trans(Geo g, Dev d, Params p) {
  if (gated(d)) return;
  const int J = blockIdx.x + 2, lane = threadIdx.x;
  const int IMM1 = g.IMM1, KMM1 = g.KMM1, IMID = p.IMID, KMID = p.KMID;
  const int NRW = d.NROW[J - 1];
  const int JROW = J - d.JSTART[NRW - 1] + 1;
  float *VR = d.VISC_RAT;
  if (g.KM != 2) {
    const int JTRHUB = d.JTRAN_K1[NRW - 1], JTRTIP = d.JTRAN_KM[NRW - 1];
    for (int I = lane + 1; I <= IMM1; I += 64) {
      bool lam = true;
      if (JROW >= JTRHUB) {
        float RMAX = 0.0f;
        for (int K = 1; K <= KMID; K++) {
          const float v = VR[IX(I, J, K)];
          if (!(v < RMAX)) RMAX = v;
        }
        lam = !(RMAX > p.FTRANS);
      }
      if (lam)
        for (int K = 1; K <= KMID; K++) VR[IX(I, J, K)] = 1.0f;
      lam = true;
      if (JROW >= JTRTIP) {
        float RMAX = 0.0f;
        for (int K = KMID; K <= KMM1; K++) {
          const float v = VR[IX(I, J, K)];
          if (!(v < RMAX)) RMAX = v;
        }
        lam = !(RMAX > p.FTRANS);
      }
      if (lam)
        for (int K = KMID; K <= KMM1; K++) VR[IX(I, J, K)] = 1.0f;
    }
  }
  __syncthreads();
  const int JTRLOW = d.JTRAN_I1[NRW - 1], JTRUP = d.JTRAN_IM[NRW - 1];
  for (int K = lane + 1; K <= KMM1; K += 64) {
    bool lam = true;
    if (JROW >= JTRLOW) {
      float RMAX = 0.0f;
      for (int I = 1; I <= IMID; I++) {
        const float v = VR[IX(I, J, K)];
        if (!(v < RMAX)) RMAX = v;
      }
      lam = !(RMAX > p.FTRANS);
    }
    if (lam)
      for (int I = 1; I <= IMID; I++) VR[IX(I, J, K)] = 1.0f;
    lam = true;
    if (JROW >= JTRUP) {
      float RMAX = 0.0f;
      for (int I = IMID; I <= IMM1; I++) {
        const float v = VR[IX(I, J, K)];
        if (!(v < RMAX)) RMAX = v;
      }
      lam = !(RMAX > p.FTRANS);
    }
    if (lam)
      for (int I = IMID; I <= IMM1; I++) VR[IX(I, J, K)] = 1.0f;
  }
}

// SPAL_LOSS's helicity and pressure-gradient source factors,
// mf:13725-13780.  The reference forms them as scalars in its cell loop and
// its source loop applies the values left by the loop's last cell
// (IM-1, JM, KM-1) to every cell (mf:13954-13955), so one thread forms that
// cell's.  The pressure-gradient factor scales with the Reynolds number of
// the last row's trailing edge, which the cell loop forms on its way
// (mf:13795-13806), and uses the laminar viscosity the loop holds on
// reaching the cell: the previous cell's when it varies with temperature.
__global__ void k_sa_factors(Geo g, Dev d, Params p) {
  if (gated(d)) return;
  if (threadIdx.x != 0) return;
  const int I = g.IMM1, J = g.JM, K = g.KMM1, IP1 = I + 1, JM1 = J - 1, KP1 = K + 1;
  const unsigned n = IX(I, J, K);
  const float *VX = d.VX, *VR = d.VR, *WT = d.WT;
  const float WTSUM = WT[IX(I, J, K)] + WT[IX(IP1, J, K)] + WT[IX(IP1, J, KP1)] + WT[IX(I, J, KP1)] +
                      WT[IX(I, JM1, K)] + WT[IX(IP1, JM1, K)] + WT[IX(IP1, JM1, KP1)] + WT[IX(I, JM1, KP1)];
  const float VXSUM = VX[IX(I, J, K)] + VX[IX(IP1, J, K)] + VX[IX(IP1, J, KP1)] + VX[IX(I, J, KP1)] +
                      VX[IX(I, JM1, K)] + VX[IX(IP1, JM1, K)] + VX[IX(IP1, JM1, KP1)] + VX[IX(I, JM1, KP1)];
  const float VRSUM = VR[IX(I, J, K)] + VR[IX(IP1, J, K)] + VR[IX(IP1, J, KP1)] + VR[IX(I, J, KP1)] +
                      VR[IX(I, JM1, K)] + VR[IX(IP1, JM1, K)] + VR[IX(IP1, JM1, KP1)] + VR[IX(I, JM1, KP1)];
  float WREL = sqrtf(VXSUM * VXSUM + VRSUM * VRSUM + WTSUM * WTSUM);
  if (p.FAC_VORT > 0.001f) {
    const float WTAVG = 0.125f * WTSUM;
    const float RAVG = d.RAVG_CELL[JKX(J, K)];
    float DVXDX, DVXDR, DVXDT, DVRDX, DVRDR, DVRDT, DWTDX, DWTDR, DWTDT;
    sa_gradvel(g, d, d.VX, I, J, K, DVXDX, DVXDR, DVXDT);
    sa_gradvel(g, d, d.VR, I, J, K, DVRDX, DVRDR, DVRDT);
    sa_gradvel(g, d, d.WT, I, J, K, DWTDX, DWTDR, DWTDT);
    const float VORTX = DVRDT - DWTDR - WTAVG / RAVG;
    const float VORTR = DWTDX - DVXDT;
    const float VORTT = DVXDR - DVRDX;
    const float HELFAC = (VXSUM * VORTX + VRSUM * VORTR + WTSUM * VORTT) / (WREL * d.ABS_VORT[n]);
    d.SC[SC_VORTFAC] = (1.0f + p.FAC_VORT * tanh_r(3.0f * HELFAC * HELFAC));
  }
  if (p.FAC_PGRAD > 0.001f) {
    const int NRW = d.NROW[J - 1], JTE = d.JTE[NRW - 1];
    float REYNOLDS = 0.0f;
    if (JTE >= 2 && JTE <= g.JM && p.IMID <= g.IMM1 && p.KMID <= g.KMM1) {
      const unsigned m = IX(p.IMID, JTE, p.KMID);
      const float WABS = sqrtf(VX[m] * VX[m] + VR[m] * VR[m] + WT[m] * WT[m]);
      const float ROVEXIT = d.RO[m] * WABS;
      REYNOLDS = d.CHORD[NRW - 1] * ROVEXIT / d.VISC_LAM[m];
    }
    const unsigned c = IX(p.IMID, J, K);
    const float WLIM = 0.1f * sqrtf(VX[c] * VX[c] + VR[c] * VR[c] + WT[c] * WT[c]);
    if (WREL < WLIM) WREL = WLIM;
    WREL = 0.125f * WREL;
    float PGRAD = d.DPDS_CELL[n];
    if (PGRAD < 0.0f) PGRAD = 0.0f;
    const float VISLAM = (p.REYNO < 0.0001f) ? d.VISC_LAM[IX(I - 1, J, K)] : d.SC[SC_VISLAM];
    const float ROAVG = d.ROAVG_CELL[n];
    float PTERM = PGRAD * VISLAM / (ROAVG * ROAVG * WREL * WREL * WREL);
    PTERM = PTERM * REYNOLDS;
    d.SC[SC_PGRADFAC] = (1.0f + p.FAC_PGRAD * tanh_r(PTERM * PTERM));
  }
}

// gradient of the kinematic turbulent viscosity, mf:13893-13911 (J <= JM-1;
// the J = JM layer stays zero as in the reference)
__global__ void k_sa_gradtvis(Geo g, Dev d, Params p) {
  if (gated(d)) return;
  CELL_IJK
  if (J > g.JMM1) return;
  const unsigned n = IX(I, J, K);
  float DX_, DR_, DT_;
  sa_gradcell(g, d, d.TRANS_KVIS, I, J, K, p.IBOUND, DX_, DR_, DT_);
  const int NRW = d.NROW[J - 1];
  if (J == d.JSTART[NRW - 1] || J == d.JSTART[NRW - 1] - 1) {
    d.GRAD_TVIS[n] = 0.0f;
    d.TVGX[n] = 0.0f;
    d.TVGR[n] = 0.0f;
    d.TVGT[n] = 0.0f;
    return;
  }
  d.GRAD_TVIS[n] = DX_ * DX_ + DR_ * DR_ + DT_ * DT_;
  const float VISCEFF = d.TRANS_KVIS[n] + d.VISC_LAM[n];
  d.TVGX[n] = VISCEFF * DX_;
  d.TVGR[n] = VISCEFF * DR_;
  d.TVGT[n] = VISCEFF * DT_;
}

// Transition tests, mf:14004-14111: per (I,J) column (hub, casing) and per
// (J,K) line (I = 1 and I = IM surfaces); bit 0/1 = damp the source there.
__global__ void k_sa_trcol(Geo g, Dev d, Params p) {
  if (gated(d)) return;
  const int I = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int J = blockIdx.y + 2;
  if (I > g.IMM1 || J > g.JM) return;
  const int KEND1 = p.KMID / 2, KENDM = g.KMM1 - KEND1;
  const int NRW = d.NROW[J - 1], JROW = J - d.JSTART[NRW - 1] + 1;
  const int JTRHUB = d.JTRAN_K1[NRW - 1], JTRTIP = d.JTRAN_KM[NRW - 1];
  int flags = 0;
  if (!(JTRHUB <= 1 && p.FTRANS < 0.1f)) {
    bool damp = JROW < JTRHUB;
    if (!damp) {
      float RMAX = 0.0f;
      for (int K = 1; K <= KEND1; K++) {
        const float r = d.VISC_RAT[IX(I, J, K)];
        if (!(r < RMAX)) RMAX = r;
      }
      damp = !(RMAX > p.FTRANS);
    }
    if (damp) flags |= 1;
  }
  if (!(JTRTIP <= 1 && p.FTRANS < 0.1f)) {
    bool damp = JROW < JTRTIP;
    if (!damp) {
      float RMAX = 0.0f;
      for (int K = KENDM; K <= g.KMM1; K++) {
        const float r = d.VISC_RAT[IX(I, J, K)];
        if (!(r < RMAX)) RMAX = r;
      }
      damp = !(RMAX > p.FTRANS);
    }
    if (damp) flags |= 2;
  }
  d.SAFLAG[(I - 1) + g.IM * (J - 1)] = flags;
}

__global__ void k_sa_trline(Geo g, Dev d, Params p) {
  if (gated(d)) return;
  const int K = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int J = blockIdx.y + 2;
  if (K > g.KMM1 || J > g.JM) return;
  const int IEND1 = p.IMID / 2, IENDM = g.IMM1 - IEND1;
  const int NRW = d.NROW[J - 1], JROW = J - d.JSTART[NRW - 1] + 1, JLEROW = d.JLE[NRW - 1];
  const int JTRLOW = d.JTRAN_I1[NRW - 1], JTRUP = d.JTRAN_IM[NRW - 1];
  int flags = 0;
  if (!(JTRLOW <= 1 && p.FTRANS < 0.1f) && !(J <= JLEROW)) {
    bool damp = JROW < JTRLOW || J < JLEROW;
    if (!damp) {
      float RMAX = 0.0f;
      for (int I = 1; I <= IEND1; I++) {
        const float r = d.VISC_RAT[IX(I, J, K)];
        if (!(r < RMAX)) RMAX = r;
      }
      damp = !(RMAX > p.FTRANS);
    }
    if (damp) flags |= 1;
  }
  if (!(JTRUP <= 1 && p.FTRANS < 0.1f) && !(J <= JLEROW)) {
    bool damp = JROW < JTRUP;
    if (!damp) {
      float RMAX = 0.0f;
      for (int I = IENDM; I <= g.IMM1; I++) {
        const float r = d.VISC_RAT[IX(I, J, K)];
        if (!(r < RMAX)) RMAX = r;
      }
      damp = !(RMAX > p.FTRANS);
    }
    if (damp) flags |= 2;
  }
  d.SAFLAG[g.IM * g.JM + JKX(J, K)] = flags;
}

// source of the turbulent viscosity, relaxed into T_SOURCE, then damped
// towards laminar where the tests say so, mf:13916-13997 and 14004-14116
__global__ void k_sa_source(Geo g, Dev d, Params p, float RF_VIS) {
  if (gated(d)) return;
  CELL_IJK
  const unsigned n = IX(I, J, K);
  const float CW2 = 0.3f, SIGMA = 0.6667f, CB2 = 0.622f, CB1 = 0.1355f, CW1 = 0.80607f + 2.433f;
  const float FST0 = p.FAC_ST0 * CB1, FST1 = p.FAC_ST1 / SIGMA, FST2 = p.FAC_ST2 * CB2 / SIGMA;
  const float FST3 = p.FAC_ST3 * CW1, FSTMIX = p.FAC_STMIX / 100000.0f;
  const float RF_VIS1 = 1.0f - RF_VIS;
  const float yp = d.Y_PLUS[n];
  float FYPLUS = 0.0f;
  if (yp > p.YPLAM) {
    float XFAC = (yp - p.YPLAM) / (p.YPTURB - p.YPLAM);
    if (XFAC > 1.0f) XFAC = 1.0f;
    FYPLUS = XFAC * XFAC * (3.0f - 2.0f * XFAC);
  }
  float TVISDDX, TVISDDR, TVISDDT, d1, d2;
  sa_gradcell(g, d, d.TVGX, I, J, K, p.IBOUND, TVISDDX, d1, d2);
  sa_gradcell(g, d, d.TVGR, I, J, K, p.IBOUND, d1, TVISDDR, d2);
  sa_gradcell(g, d, d.TVGT, I, J, K, p.IBOUND, d1, d2, TVISDDT);
  const float ROAVG = d.ROAVG_CELL[n];
  const float TVIS = d.TRANS_KVIS[n];
  const float TV = d.ABS_VORT[n];
  float ST0 = FST0 * TV * TVIS;
  if (p.FAC_VORT > 0.001f) ST0 = ST0 * d.SC[SC_VORTFAC];
  if (p.FAC_PGRAD > 0.001f) ST0 = ST0 * d.SC[SC_PGRADFAC];
  ST0 = ST0 * FYPLUS;
  const float ST1 = FST1 * (TVISDDX + TVISDDR + TVISDDT);
  const float ST2 = FST2 * d.GRAD_TVIS[n];
  const float DW = d.DWALLSQ[n];
  float TAU = 5.9488f * TVIS / TV / DW;
  if (TAU > 1.2f) TAU = 1.2f;
  if (TAU < 0.001f) TAU = 0.001f;
  const float G = TAU + CW2 * (TAU * TAU * TAU * TAU * TAU * TAU - TAU);
  const float G6CW6 = G * G * G * G * G * G / 64.0f;
  const float FW = G * pow_r((1 + 0.015625f) / (1 + G6CW6), 0.166667f);
  const float ST3 = FST3 * FW * TVIS * TVIS / DW;
  const float ST_MIXL = FSTMIX * (d.DYN_VIS_ML[n] - d.VISC_TURB[n]) / d.STEP[n];
  float ts = RF_VIS1 * d.T_SOURCE[n] + RF_VIS * (ROAVG * (ST0 + ST1 + ST2 - ST3) * d.VOL[n] + ST_MIXL);
  if (J > 2 && d.INDMIX[J - 2] == 1) ts = 0.0f;
  if (J > 2 && d.INDMIX[J - 3] == 1) ts = 0.0f;
  // transition: hub and casing first, then the blade surfaces (mf:14004-14116)
  const int KEND1 = p.KMID / 2, KENDM = g.KMM1 - KEND1, IEND1 = p.IMID / 2, IENDM = g.IMM1 - IEND1;
  const float VT = d.VISC_TURB[n], ST = d.STEP[n];
  const int fc = d.SAFLAG[(I - 1) + g.IM * (J - 1)], fl = d.SAFLAG[g.IM * g.JM + JKX(J, K)];
  if ((fc & 1) && K <= KEND1) {
    const float TVEDGE = d.VISC_TURB[IX(I, J, KEND1)];
    float FK = (float)(KEND1 + 1 - K) / (float)KEND1;
    FK = FK * FK * (3 - 2 * FK);
    const float TVSPEC = FK * TVEDGE;
    ts = ts - (VT - TVSPEC) / ST / 100000.0f;
  }
  if ((fc & 2) && K >= KENDM) {
    const float TVEDGE = d.VISC_TURB[IX(I, J, KENDM)];
    float FK = (float)(K + 1 - KENDM) / (float)(g.KM - KENDM);
    FK = FK * FK * (3 - 2 * FK);
    const float TVSPEC = FK * TVEDGE;
    ts = ts - (VT - TVSPEC) / ST / 100000.0f;
  }
  if ((fl & 1) && I <= IEND1) {
    const float TVEDGE = d.VISC_TURB[IX(IEND1, J, K)];
    float FI = (float)(IEND1 + 1 - I) / (float)IEND1;
    FI = FI * FI * (3 - 2 * FI);
    const float TVSPEC = FI * TVEDGE;
    ts = ts - (VT - TVSPEC) / ST / 100000.0f;
  }
  if ((fl & 2) && I >= IENDM) {
    const float TVEDGE = d.VISC_TURB[IX(IENDM, J, K)];
    float FI = (float)(I + 1 - IENDM) / (float)(g.IM - IENDM);
    FI = FI * FI * (3 - 2 * FI);
    const float TVSPEC = FI * TVEDGE;
    ts = ts - (VT - TVSPEC) / ST / 100000.0f;
  }
  d.T_SOURCE[n] = ts;
}

// wall friction of SPAL_LOSS's wall cells (sums of four node values: the
// 1/16 and 1/128 factors), mf:14221-14268
__device__ float sa_cf(float RE, float DPERP, float ROUGH, float ROAVG4, float WAVG, float VISLAM) {
  const float RELOG = 1.f / log_r(RE);
  float CF;
  if (ROUGH > 1.0e-7f) {
    if (ROUGH > DPERP) ROUGH = DPERP;
    float REK = RE * ROUGH / DPERP;
    REK = REK - 80.0f;
    if (REK < 0.0f) REK = 0.0f;
    const float REKSQ = REK * REK;
    const float A1 = -.00178493f + .0000814923f * REK + .000000150445f * REKSQ;
    const float A2 = .029072f - .001584f * REK - .00000225194f * REKSQ;
    const float A3 = .270313f + .0091409f * REK + .00000451537f * REKSQ + .00000000464767f * REKSQ * REK;
    CF = A1 + A2 * RELOG + A3 * RELOG * RELOG;
  } else {
    CF = -0.00178493f + 0.029072f * RELOG + 0.270313f * RELOG * RELOG;
  }
  const float CFLAM = 2.0f / RE;
  CF = fmaxf(CF, CFLAM);
  if (RE < 125.f) CF = CFLAM;
  if (ROUGH > 1.0e-7f) {
    const float TW = CF * ROAVG4 * WAVG * WAVG / 128.f;
    const float VSTAR = sqrtf(TW / (0.25f * ROAVG4));
    const float PLUSK = ROUGH * VSTAR * ROAVG4 / VISLAM / 4.f;
    const float RELOG10 = log10f(DPERP / ROUGH);
    const float FUNCN = 5.75f * RELOG10 + 8.5f;
    const float CF_FULL_ROUGH = 2.0f / (FUNCN * FUNCN);
    if (PLUSK > 45.f && CF > CF_FULL_ROUGH) CF = CF_FULL_ROUGH;
  }
  return CF;
}

// the nine stresses and heat fluxes of a cell with its final viscosity,
// mf:14123-14147
struct SaStress {
  float xr, xx, xt, rr, rt, tt, qx, qr, qt;
};

__device__ __forceinline__ SaStress sa_stress(const Geo &g, const Dev &d, const Params &p, unsigned n) {
  const float VISLAM = d.VISC_LAM[n];
  const float VISTOT = d.VISC_RAT[n] * VISLAM;
  const float THCOND = p.CP * VISTOT / p.PRANDTL;
  const float VISTOT2 = 2.0f * VISTOT;
  SaStress s;
  s.xr = VISTOT * d.SXR[n];
  s.xx = VISTOT2 * d.SXX[n];
  s.xt = VISTOT * d.SXT[n];
  s.rr = VISTOT2 * d.SRR[n];
  s.rt = VISTOT * d.SRT[n];
  s.tt = VISTOT2 * d.STT[n];
  s.qx = -THCOND * d.GQX[n];
  s.qr = -THCOND * d.GQR[n];
  s.qt = -THCOND * d.GQT[n];
  return s;
}

// Forces from the blade surfaces and the I = const faces, mf:14159-14500.
// One wavefront per (J,K) line.  d.SAJMAP gives, per J, the J-1 and the row
// the reference's periodic-force block reads (its loop variables keep the
// values of the last blade J visited: [0..JM) and [JM..2JM) the J-1 for
// K = 1 and K > 1, [2JM..4JM) the row likewise).
constexpr int kSaLine = 260;

__global__ void __launch_bounds__(64) k_sa_iforces(Geo g, Dev d, Params p) {
  if (gated(d)) return;
  __shared__ float TXR[kSaLine], TXT[kSaLine], TXX[kSaLine], TRR[kSaLine], TRT[kSaLine], TTT[kSaLine];
  __shared__ float QXX[kSaLine], QRR[kSaLine], QTT[kSaLine];
  __shared__ float XF[kSaLine], RF[kSaLine], TF[kSaLine], QF[kSaLine], WF[kSaLine];
  __shared__ float F0[4], F1[4], VS[2];
  const int K = blockIdx.x + 1, J = blockIdx.y + 2, lane = threadIdx.x;
  const int IM = g.IM, IMM1 = g.IMM1, IMID = p.IMID, KP1 = K + 1, JM = g.JM;
  const float VISLAM = d.VISC_LAM[IX(IMM1, JM, g.KMM1)], VISLAM16 = VISLAM * 16.0f;
  const float ABTJK = d.ABT[JKX(J, K)];
  for (int I = lane + 1; I <= IMM1; I += 64) {
    const SaStress s = sa_stress(g, d, p, IX(I, J, K));
    TXR[I] = s.xr; TXX[I] = s.xx; TXT[I] = s.xt; TRR[I] = s.rr; TRT[I] = s.rt; TTT[I] = s.tt;
    QXX[I] = s.qx; QRR[I] = s.qr; QTT[I] = s.qt;
  }
  if (lane < 4) { F0[lane] = 0.0f; F1[lane] = 0.0f; }
  __syncthreads();
  // blade-surface friction and y+ (mf:14159-14424)
  const bool blade = d.IND[J - 1] != 0;
  const int NRWJ = d.NROW[J - 1];
  const bool tipcell = (d.KTIPS[NRWJ - 1] > 0) && (K >= d.KTIPS[NRWJ - 1]) && (K < d.KTIPE[NRWJ - 1]);
  if (blade && !tipcell) {
    // the I = 1 wall on lane 0 and the I = IM wall on lane 1 (mf:14189-14394)
    if (lane < 2) {
      const bool lo = lane == 0;
      const int In = lo ? 2 : IMM1, Ic = lo ? 1 : IMM1, Ia = lo ? 1 : IM, JM1 = J - 1;
      const float *VX = d.VX, *VR = d.VR, *WT = d.WT, *RO = d.RO;
      const float VXAVG = VX[IX(In, J, K)] + VX[IX(In, JM1, K)] + VX[IX(In, JM1, KP1)] + VX[IX(In, J, KP1)];
      const float VRAVG = VR[IX(In, J, K)] + VR[IX(In, JM1, K)] + VR[IX(In, JM1, KP1)] + VR[IX(In, J, KP1)];
      const float WTAVG = WT[IX(In, J, K)] + WT[IX(In, JM1, K)] + WT[IX(In, JM1, KP1)] + WT[IX(In, J, KP1)];
      const float ROAVG = RO[IX(In, J, K)] + RO[IX(In, JM1, K)] + RO[IX(In, JM1, KP1)] + RO[IX(In, J, KP1)];
      const float WAVG = sqrtf(WTAVG * WTAVG + VXAVG * VXAVG + VRAVG * VRAVG);
      const float bx = d.ABX[IX(Ia, J, K)], br = d.ABR[IX(Ia, J, K)];
      const float AREA = sqrtf(bx * bx + br * br + ABTJK * ABTJK);
      const float DPERP = d.VOL[IX(Ic, J, K)] / AREA;
      float TWALL;
      if (p.YPLUSWALL < -0.001f) {  // Shih wall functions, mf:14209-14219, 14326-14336
        const unsigned c = IX(Ic, J, K);
        TWALL = wallfun(p.YPLUSWALL, VISLAM, DPERP, d.DPDS_CELL[c], &d.DPDS_CELL[c], 0.25f * ROAVG,
                        &(lo ? d.YPLUS_I1 : d.YPLUS_IM)[JKX(J, K)], 0.25f * WAVG, true);
      } else {
        const float RE = DPERP * ROAVG * WAVG / VISLAM16;
        const float CF = sa_cf(RE, DPERP, (lo ? d.ROUGH_L : d.ROUGH_U)[NRWJ - 1], ROAVG, WAVG, VISLAM);
        TWALL = CF * ROAVG * WAVG * WAVG / 128.f;
      }
      VS[lane] = sqrtf(TWALL / (0.25f * ROAVG));
      float *F = lo ? F0 : F1;
      F[3] = ROAVG;
      const float FMULT = -TWALL * AREA / WAVG;
      const float TFORCE = FMULT * WTAVG;
      F[0] = 0.0f + FMULT * VXAVG;
      F[1] = 0.0f + FMULT * VRAVG;
      F[2] = 0.0f + TFORCE;
      // energy of the wall cells in XF slots 0 and IM (not faces)
      XF[lo ? 0 : IM] = TFORCE * d.WRAD[J - 1] * d.RAVG_CELL[JKX(J, K)];
    }
    __syncthreads();
    // SPAL_LOSS sets y+ up to I = IM (mf:14404), NEW_LOSS up to IM-1 (mf:13035)
    const int iend = (p.ILOS >= 200) ? IM : IMM1;
    for (int I = lane + 1; I <= iend; I += 64) {
      const unsigned n = IX(I, J, K);
      const float YPLS = (I <= IMID) ? VS[0] * sqrtf(d.DWALLSQ[n]) * (0.25f * F0[3]) / VISLAM
                                     : VS[1] * sqrtf(d.DWALLSQ[n]) * (0.25f * F1[3]) / VISLAM;
      d.Y_PLUS[n] = fminf(1000.0f, YPLS);
    }
  }
  // running sums of cells 1 and IM-1 before the I-face loop
  float fx1 = 0.f, fr1 = 0.f, ft1 = 0.f, es1 = 0.f, fxm = 0.f, frm = 0.f, ftm = 0.f, esm = 0.f;
  if (blade && !tipcell) {
    fx1 = F0[0]; fr1 = F0[1]; ft1 = F0[2]; es1 = 0.0f + XF[0];
    fxm = 0.0f + F1[0]; frm = 0.0f + F1[1]; ftm = 0.0f + F1[2]; esm = 0.0f + XF[IM];
  }
  __syncthreads();
  // periodic cells I = 1 and IM-1, mf:14434-14486 (the reference's
  // J-1 and row here are those of the last blade J it visited)
  const int JM1E = (K == 1) ? d.SAJMAP[J - 1] : d.SAJMAP[JM + J - 1];
  const int NRWE = (K == 1) ? d.SAJMAP[2 * JM + J - 1] : d.SAJMAP[3 * JM + J - 1];
  const bool skipper = ((d.IND[J - 1] == 1) && (K < d.KTIPS[NRWE - 1])) || (K > d.KTIPE[NRWE - 1]);
  if (!skipper) {
    const int I = 1, IM1 = IMM1;
    const float abr = d.ABR[IX(I, J, K)], abx = d.ABX[IX(I, J, K)];
    const float XFOR = 0.5f * ((TXR[I] + TXR[IM1]) * abr + (TXT[I] + TXT[IM1]) * ABTJK + (TXX[I] + TXX[IM1]) * abx);
    fxm = fxm + XFOR;
    fx1 = fx1 - XFOR;
    float AVG = 0.5f * (fx1 + fxm);
    fx1 = AVG; fxm = AVG;
    const float RFOR = 0.5f * ((TRR[I] + TRR[IM1]) * abr + (TRT[I] + TRT[IM1]) * ABTJK + (TXR[I] + TXR[IM1]) * abx);
    frm = frm + RFOR;
    fr1 = fr1 - RFOR;
    AVG = 0.5f * (fr1 + frm);
    fr1 = AVG; frm = AVG;
    const float TFOR = 0.5f * ((TRT[I] + TRT[IM1]) * abr + (TTT[I] + TTT[IM1]) * ABTJK + (TXT[I] + TXT[IM1]) * abx);
    ftm = ftm + TFOR;
    ft1 = ft1 - TFOR;
    AVG = 0.5f * (ft1 + ftm);
    ft1 = AVG; ftm = AVG;
    const float QFLOW = 0.5f * ((QXX[I] + QXX[IM1]) * abx + (QRR[I] + QRR[IM1]) * abr + (QTT[I] + QTT[IM1]) * ABTJK);
    const float *VX = d.VX, *VR = d.VR, *VT = d.VT;
    const float VXAVG = VX[IX(I, J, K)] + VX[IX(I, JM1E, K)] + VX[IX(I, JM1E, KP1)] + VX[IX(I, J, KP1)];
    const float VRAVG = VR[IX(I, J, K)] + VR[IX(I, JM1E, K)] + VR[IX(I, JM1E, KP1)] + VR[IX(I, J, KP1)];
    const float VTAVG = VT[IX(I, J, K)] + VT[IX(I, JM1E, K)] + VT[IX(I, JM1E, KP1)] + VT[IX(I, J, KP1)];
    const float WVISC = -0.25f * (XFOR * VXAVG + RFOR * VRAVG + TFOR * VTAVG);
    es1 = es1 + QFLOW + WVISC;
    esm = esm - QFLOW - WVISC;
    AVG = 0.5f * (es1 + esm);
    es1 = AVG; esm = AVG;
  }
  // the I = const faces, mf:14491-14498
  for (int I = lane + 2; I <= IMM1; I += 64) {
    const int IM1 = I - 1;
    const unsigned n = IX(I, J, K);
    const float abr = d.ABR[n], abx = d.ABX[n];
    const float XFOR = 0.5f * ((TXR[I] + TXR[IM1]) * abr + (TXT[I] + TXT[IM1]) * ABTJK + (TXX[I] + TXX[IM1]) * abx);
    const float RFOR = 0.5f * ((TRR[I] + TRR[IM1]) * abr + (TRT[I] + TRT[IM1]) * ABTJK + (TXR[I] + TXR[IM1]) * abx);
    const float TFOR = 0.5f * ((TRT[I] + TRT[IM1]) * abr + (TTT[I] + TTT[IM1]) * ABTJK + (TXT[I] + TXT[IM1]) * abx);
    const float QFLOW = 0.5f * ((QXX[I] + QXX[IM1]) * abx + (QRR[I] + QRR[IM1]) * abr + (QTT[I] + QTT[IM1]) * ABTJK);
    const float *VX = d.VX, *VR = d.VR, *VT = d.VT;
    const float VXAVG = VX[n] + VX[IX(I, JM1E, K)] + VX[IX(I, JM1E, KP1)] + VX[IX(I, J, KP1)];
    const float VRAVG = VR[n] + VR[IX(I, JM1E, K)] + VR[IX(I, JM1E, KP1)] + VR[IX(I, J, KP1)];
    const float VTAVG = VT[n] + VT[IX(I, JM1E, K)] + VT[IX(I, JM1E, KP1)] + VT[IX(I, J, KP1)];
    XF[I] = XFOR;
    RF[I] = RFOR;
    TF[I] = TFOR;
    QF[I] = QFLOW;
    WF[I] = -0.25f * (XFOR * VXAVG + RFOR * VRAVG + TFOR * VTAVG);
  }
  __syncthreads();
  // each cell: its start value, minus its own face, plus the next face (the
  // reference's loop adds in this order)
  for (int c = lane + 1; c <= IMM1; c += 64) {
    float fx = 0.f, fr = 0.f, ft = 0.f, es = 0.f;
    if (c == 1) { fx = fx1; fr = fr1; ft = ft1; es = es1; }
    if (c == IMM1) { fx = fxm; fr = frm; ft = ftm; es = esm; }
    if (c >= 2) {
      fx = fx - XF[c];
      fr = fr - RF[c];
      ft = ft - TF[c];
      es = es + QF[c] + WF[c];
    }
    if (c + 1 <= IMM1) {
      fx = fx + XF[c + 1];
      fr = fr + RF[c + 1];
      ft = ft + TF[c + 1];
      es = es - QF[c + 1] - WF[c + 1];
    }
    const unsigned n = IX(c, J, K);
    d.FORCEX[n] = fx;
    d.FORCER[n] = fr;
    d.FORCET[n] = ft;
    d.ESOURCE[n] = es;
  }
}

// Hub and casing friction of each (I,J) column, mf:14505-14766 (SPAL_LOSS;
// NEW_LOSS mf:13167-13400): the forces and energy source the two wall
// cells receive, and the friction velocity and density the column's y+ is
// formed with, into d.KWALL (kKWall quantities, each a plane of
// (IM-1) x (JM-1) columns).  One thread per column.
__global__ void k_sa_kwalls(Geo g, Dev d, Params p) {
  if (gated(d)) return;
  const int I = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int J = blockIdx.y * blockDim.y + threadIdx.y + 2;
  if (I > g.IMM1 || J > g.JM) return;
  const int IP1 = I + 1, JM1 = J - 1, KMM1 = g.KMM1, KM = g.KM;
  const float VISLAM = d.VISC_LAM[IX(g.IMM1, g.JM, KMM1)], VISLAM16 = VISLAM * 16.0f;
  // SPAL_LOSS's hub/casing loop reads the row its blade loop left
  // (mf:14536); NEW_LOSS's takes the row of its own J (mf:13169)
  const int NRWHC = (p.ILOS >= 200) ? d.SAJMAP[4 * g.JM] : d.NROW[J - 1];
  const float *VX = d.VX, *VR = d.VR, *VT = d.VT, *RO = d.RO;
  // hub, mf:14510-14650
  float VXAVG1 = VX[IX(I, J, 2)] + VX[IX(IP1, J, 2)] + VX[IX(IP1, JM1, 2)] + VX[IX(I, JM1, 2)];
  float VRAVG1 = VR[IX(I, J, 2)] + VR[IX(IP1, J, 2)] + VR[IX(IP1, JM1, 2)] + VR[IX(I, JM1, 2)];
  float VTAVG1 = VT[IX(I, J, 2)] + VT[IX(IP1, J, 2)] + VT[IX(IP1, JM1, 2)] + VT[IX(I, JM1, 2)];
  float ROAVG1 = RO[IX(I, J, 2)] + RO[IX(IP1, J, 2)] + RO[IX(IP1, JM1, 2)] + RO[IX(I, JM1, 2)];
  float RAVG1 = d.R[JKX(J, 2)] + d.R[JKX(J, 2)] + d.R[JKX(JM1, 2)] + d.R[JKX(JM1, 2)];
  float WTAVG1 = VTAVG1 - d.WHUB[J - 1] * RAVG1;
  float WAVG = sqrtf(VXAVG1 * VXAVG1 + VRAVG1 * VRAVG1 + WTAVG1 * WTAVG1);
  float sx = d.ASX[IX(I, J, 1)], sr = d.ASR[IX(I, J, 1)];
  const float AREA1 = sqrtf(sx * sx + sr * sr);
  float DPERP = d.VOL[IX(I, J, 1)] / AREA1;
  float TWALLK1;
  if (p.YPLUSWALL < -0.001f) {  // Shih wall functions, mf:14556-14566
    const unsigned c1 = IX(I, J, 1);
    TWALLK1 = wallfun(p.YPLUSWALL, VISLAM, DPERP, d.DPDS_CELL[c1], &d.DPDS_CELL[c1], 0.25f * ROAVG1,
                      &d.YPLUS_K1[IJX(I, J)], 0.25f * WAVG, true);
  } else {
    const float RE = DPERP * ROAVG1 * WAVG / VISLAM16;
    const float CF = sa_cf(RE, DPERP, d.ROUGH_H[NRWHC - 1], ROAVG1, WAVG, VISLAM);
    TWALLK1 = CF * ROAVG1 * WAVG * WAVG / 128.f;
  }
  float FMULT = -TWALLK1 * AREA1 / WAVG;
  if (p.IBOUND == 1 || p.IBOUND > 2) FMULT = 0.0f;
  const float XFORCE1 = FMULT * VXAVG1, RFORCE1 = FMULT * VRAVG1, TFORCE1 = FMULT * WTAVG1;
  const float WVISC1 = TFORCE1 * 0.25f * d.WHUB[J - 1] * RAVG1;
  // casing, mf:14652-14766
  float VXAVGKM = VX[IX(I, J, KMM1)] + VX[IX(IP1, J, KMM1)] + VX[IX(IP1, JM1, KMM1)] + VX[IX(I, JM1, KMM1)];
  float VRAVGKM = VR[IX(I, J, KMM1)] + VR[IX(IP1, J, KMM1)] + VR[IX(IP1, JM1, KMM1)] + VR[IX(I, JM1, KMM1)];
  float VTAVGKM = VT[IX(I, J, KMM1)] + VT[IX(IP1, J, KMM1)] + VT[IX(IP1, JM1, KMM1)] + VT[IX(I, JM1, KMM1)];
  float ROAVGKM = RO[IX(I, J, KMM1)] + RO[IX(IP1, J, KMM1)] + RO[IX(IP1, JM1, KMM1)] + RO[IX(I, JM1, KMM1)];
  float RAVGKM = d.R[JKX(J, KMM1)] + d.R[JKX(J, KMM1)] + d.R[JKX(JM1, KMM1)] + d.R[JKX(JM1, KMM1)];
  float WTAVGKM = VTAVGKM - d.WTIP[J - 1] * RAVGKM;
  WAVG = sqrtf(VXAVGKM * VXAVGKM + VRAVGKM * VRAVGKM + WTAVGKM * WTAVGKM);
  sx = d.ASX[IX(I, J, KM)];
  sr = d.ASR[IX(I, J, KM)];
  const float AREAKM = sqrtf(sx * sx + sr * sr);
  DPERP = d.VOL[IX(I, J, KMM1)] / AREAKM;
  float TWALLKM;
  if (p.YPLUSWALL < -0.001f) {  // mf:14671-14681
    const unsigned cm = IX(I, J, KMM1);
    TWALLKM = wallfun(p.YPLUSWALL, VISLAM, DPERP, d.DPDS_CELL[cm], &d.DPDS_CELL[cm], 0.25f * ROAVGKM,
                      &d.YPLUS_KM[IJX(I, J)], 0.25f * WAVG, true);
  } else {
    const float RE = DPERP * ROAVGKM * WAVG / VISLAM16;
    const float CF = sa_cf(RE, DPERP, d.ROUGH_T[NRWHC - 1], ROAVGKM, WAVG, VISLAM);
    TWALLKM = CF * ROAVGKM * WAVG * WAVG / 128.f;
  }
  FMULT = -TWALLKM * AREAKM / WAVG;
  if (p.IBOUND >= 2) FMULT = 0.0f;
  const float XFORCEKM = FMULT * VXAVGKM, RFORCEKM = FMULT * VRAVGKM, TFORCEKM = FMULT * WTAVGKM;
  const float WVISCKM = TFORCEKM * 0.25f * d.WTIP[J - 1] * RAVGKM;
  const size_t ncol = (size_t)g.IMM1 * (g.JM - 1), col = (size_t)(I - 1) + (size_t)g.IMM1 * (J - 2);
  float *W = d.KWALL;
  W[0 * ncol + col] = XFORCE1;
  W[1 * ncol + col] = RFORCE1;
  W[2 * ncol + col] = TFORCE1;
  W[3 * ncol + col] = WVISC1;
  W[4 * ncol + col] = sqrtf(TWALLK1 / (0.25f * ROAVG1));
  W[5 * ncol + col] = ROAVG1;
  W[6 * ncol + col] = XFORCEKM;
  W[7 * ncol + col] = RFORCEKM;
  W[8 * ncol + col] = TFORCEKM;
  W[9 * ncol + col] = WVISCKM;
  W[10 * ncol + col] = sqrtf(TWALLKM / (0.25f * ROAVGKM));
  W[11 * ncol + col] = ROAVGKM;
}

// The end walls' y+ (mf:14640-14648, 14756-14764), the K = const faces and
// the body forces (mf:14770-14819) for a segment of kKSeg cells of one
// (I,J) column per thread; the face below the segment is formed again from
// the cell under it.
__global__ void k_sa_kfaces(Geo g, Dev d, Params p, float RF_VIS) {
  if (gated(d)) return;
  const int I = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int J = blockIdx.y * blockDim.y + threadIdx.y + 2;
  const int k0 = blockIdx.z * kKSeg + 1;
  if (I > g.IMM1 || J > g.JM || k0 > g.KMM1) return;
  const int KMM1 = g.KMM1, KMID = p.KMID, IP1 = I + 1, JM1 = J - 1;
  const int k1 = min(k0 + kKSeg - 1, KMM1);
  const float VISLAM = d.VISC_LAM[IX(g.IMM1, g.JM, KMM1)];
  const size_t ncol = (size_t)g.IMM1 * (g.JM - 1), col = (size_t)(I - 1) + (size_t)g.IMM1 * (J - 2);
  const float *W = d.KWALL;
  const float VSTAR1 = W[4 * ncol + col], ROAVG1 = W[5 * ncol + col];
  const float VSTARKM = W[10 * ncol + col], ROAVGKM = W[11 * ncol + col];
  for (int K = k0; K <= k1; K++) {
    const unsigned n = IX(I, J, K);
    if (K <= KMID && (p.IBOUND == 0 || p.IBOUND == 2)) {
      const float YPLS = VSTAR1 * sqrtf(d.DWALLSQ[n]) * (0.25f * ROAVG1) / VISLAM;
      d.Y_PLUS[n] = fminf(YPLS, d.Y_PLUS[n]);
    }
    if (K > KMID && (p.IBOUND == 0 || p.IBOUND == 1)) {
      const float YPLS = VSTARKM * sqrtf(d.DWALLSQ[n]) * (0.25f * ROAVGKM) / VISLAM;
      d.Y_PLUS[n] = fminf(YPLS, d.Y_PLUS[n]);
    }
  }
  const float *VX = d.VX, *VR = d.VR, *VT = d.VT;
  // the face between cells K-1 and K (K >= 2) and its viscous work
  auto face = [&](int K, const SaStress &lo, const SaStress &hi, float &xf, float &rf, float &tf, float &qf,
                  float &wf) {
    const unsigned m = IX(I, J, K);
    const float asr = d.ASR[m], asx = d.ASX[m];
    xf = 0.5f * ((hi.xr + lo.xr) * asr + (hi.xx + lo.xx) * asx);
    rf = 0.5f * ((hi.rr + lo.rr) * asr + (hi.xr + lo.xr) * asx);
    tf = 0.5f * ((hi.rt + lo.rt) * asr + (hi.xt + lo.xt) * asx);
    qf = 0.5f * ((hi.qx + lo.qx) * asx + (hi.qr + lo.qr) * asr);
    const float VXAVG = VX[IX(I, J, K)] + VX[IX(IP1, J, K)] + VX[IX(IP1, JM1, K)] + VX[IX(I, JM1, K)];
    const float VRAVG = VR[IX(I, J, K)] + VR[IX(IP1, J, K)] + VR[IX(IP1, JM1, K)] + VR[IX(I, JM1, K)];
    const float VTAVG = VT[IX(I, J, K)] + VT[IX(IP1, J, K)] + VT[IX(IP1, JM1, K)] + VT[IX(I, JM1, K)];
    wf = -0.25f * (xf * VXAVG + rf * VRAVG + tf * VTAVG);
  };
  const float RF_VIS1 = 1.0f - RF_VIS;
  SaStress sb = sa_stress(g, d, p, IX(I, J, k0));
  float xfK = 0.f, rfK = 0.f, tfK = 0.f, qfK = 0.f, wfK = 0.f;  // face K (between K-1 and K)
  if (k0 >= 2) face(k0, sa_stress(g, d, p, IX(I, J, k0 - 1)), sb, xfK, rfK, tfK, qfK, wfK);
  for (int K = k0; K <= k1; K++) {
    const unsigned n = IX(I, J, K);
    float xfN = 0.f, rfN = 0.f, tfN = 0.f, qfN = 0.f, wfN = 0.f;  // face K+1
    if (K + 1 <= KMM1) {
      const SaStress st = sa_stress(g, d, p, IX(I, J, K + 1));
      face(K + 1, sb, st, xfN, rfN, tfN, qfN, wfN);
      sb = st;
    }
    float fx = d.FORCEX[n], fr = d.FORCER[n], ft = d.FORCET[n], es = d.ESOURCE[n];
    if (K == 1) {
      fx = fx + W[0 * ncol + col]; fr = fr + W[1 * ncol + col]; ft = ft + W[2 * ncol + col]; es = es + W[3 * ncol + col];
    }
    if (K == KMM1) {
      fx = fx + W[6 * ncol + col]; fr = fr + W[7 * ncol + col]; ft = ft + W[8 * ncol + col]; es = es + W[9 * ncol + col];
    }
    if (K >= 2) { fx = fx - xfK; fr = fr - rfK; ft = ft - tfK; es = es + qfK + wfK; }
    if (K + 1 <= KMM1) { fx = fx + xfN; fr = fr + rfN; ft = ft + tfN; es = es - qfN - wfN; }
    d.XFORCE[n] = RF_VIS1 * d.XFORCE[n] - RF_VIS * fx;
    d.TFORCE[n] = RF_VIS1 * d.TFORCE[n] - RF_VIS * ft * d.RAVG_CELL[JKX(J, K)];
    d.RFORCE[n] = RF_VIS1 * d.RFORCE[n] - RF_VIS * fr;
    d.QSOURCE[n] = RF_VIS1 * d.QSOURCE[n] - RF_VIS * es;
    xfK = xfN; rfK = rfN; tfK = tfN; qfK = qfN; wfK = wfN;
  }
}

// ---------------------------------------------------------------------------
// limiters, mf:5115-5168

// The bookkeeping at the top of LOOP's step (mf:3757-3824) on the device:
// the step count, the stop tests (NSTEPS_MAX; the convergence figures of the
// last check; a pending stop request on a step divisible by 10), and the
// smoothing/damping and FMIXUP ramps.  A run that has ended (IFEND set by
// the previous step, or a non-finite check) marks this step and all later
// ones as not taken; every launch of the step is gated on that.
constexpr int kFlowChecks = 10;  // checks (50 steps) the FLOW_CONV test must hold for
__global__ void k_step_begin(Geo g, Dev d, Params p) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int *C = d.CTL;
  float *S = d.SC;
  if (C[CTL_DONE] || C[CTL_IFEND]) {
    C[CTL_DONE] = 1;
    C[CTL_SKIP] = 1;
    C[CTL_NOTEND] = 1;
    return;
  }
  const int NSTEP = C[CTL_NSTEP] + 1;
  C[CTL_NSTEP] = NSTEP;
  int IFEND = 0;
  if (NSTEP == p.NSTEPS_MAX) IFEND = 1;
  if (S[SC_EAVG] < p.CONLIM && S[SC_ECONT] < 0.01f) IFEND = 1;
  if (fabsf(S[SC_EMAX]) < (4 * p.CONLIM) && S[SC_ECONT] < 0.01f) IFEND = 1;
  if (NSTEP % 10 == 0 && C[CTL_STOPREQ]) IFEND = 1;
  // the library's optional stop test on the inlet mass flow (FLOW_CONV > 0):
  // steady to FLOW_CONV over the last kFlowChecks checks, and ECONT < 0.01
  if (p.FLOW_CONV > 0.0f && C[CTL_FLOWOK] >= kFlowChecks && S[SC_ECONT] < 0.01f) IFEND = 1;
  C[CTL_IFEND] = IFEND;
  C[CTL_SKIP] = 0;
  C[CTL_NOTEND] = IFEND ? 0 : 1;
  // the smoothing and damping ramp (mf:3787-3798; from a restart file it
  // stops after step 100)
  const float FCHANGE = 1.0f - (float)(NSTEP - 1) / (float)p.NCHANGE;
  if (!(p.IF_RESTART != 0 && NSTEP > 100) && FCHANGE > 0.0f) {
    S[SC_SFT] = p.SFTIN + 0.02f * FCHANGE;
    S[SC_SFX] = p.SFXIN + 0.02f * FCHANGE;
    S[SC_DAMP] = p.DAMPIN * (1.f - 0.75f * FCHANGE);
  }
  float FMIXUP = 1.0f + (p.FACMIXUP - 1.f) * (float)(p.NMIXUP - NSTEP) / (float)p.NMIXUP;
  if (FMIXUP < 1.0f) FMIXUP = 1.0f;
  S[SC_FMIXUP] = FMIXUP * FMIXUP;
}

// After a convergence check: a non-finite maximum velocity or inlet flow
// ends the run here and records the step for the host's error report; the
// relative change of the inlet mass flow since the previous check (FP64)
// feeds the optional FLOW_CONV stop test.
__global__ void k_nan_check(Geo g, Dev d, Params p) {
  if (gated(d)) return;
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  float *S = d.SC;
  if (!isfinite(S[SC_VMAX]) || !isfinite(S[SC_FLOWRAT])) {
    d.CTL[CTL_NANSTEP] = d.CTL[CTL_NSTEP];
    d.CTL[CTL_DONE] = 1;
  }
  const double f = S[SC_FLOWRAT], f0 = S[SC_FLOWPREV];
  S[SC_DFLOW] = f != 0.0 ? (float)(fabs(f - f0) / fabs(f)) : 1.0e6f;
  S[SC_FLOWPREV] = S[SC_FLOWRAT];
  d.CTL[CTL_FLOWOK] = S[SC_DFLOW] < p.FLOW_CONV ? d.CTL[CTL_FLOWOK] + 1 : 0;
}

__global__ void k_fill_int(int *p, int n, int v) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

__device__ __forceinline__ float rolim_one(float ro, float rom2, float rom1, float rop1, float rop2, float ROLIM) {
  float ROAVG = 0.25f * (rom2 + rom1 + rop1 + rop2);
  if (ro < 0.6f * ROAVG) ro = 0.5f * (ro + 0.6f * ROAVG);
  if (ro > 1.5f * ROAVG) ro = 0.5f * (ro + 1.5f * ROAVG);
  if (ro < ROLIM) ro = 0.5f * (ro + ROLIM);
  return ro;
}

// Density limiter (mf:5115-5122) in two passes.  The reference updates RO in
// place along J, so a change at J is seen by J+1 and J+2.  Pass 1 evaluates
// every node with the unmodified field and flags the (I,K) lines where any
// node would change; pass 2 replays only those lines serially in J.
__global__ void k_rolim_check(Geo g, Dev d, StepScalars s) {
  if (gated(d)) return;
  NODE_IJK
  if (J < 3 || J > g.JM - 3) return;
  float ro = d.RO[IX(I, J, K)];
  float nv = rolim_one(ro, d.RO[IX(I, J - 2, K)], d.RO[IX(I, J - 1, K)], d.RO[IX(I, J + 1, K)], d.RO[IX(I, J + 2, K)], s.ROLIM);
  if (nv != ro) d.FLAGS[IKX(I, K)] = 1;
}

// reference relative Mach conditions of row NR, mf:5127-5145
__device__ __forceinline__ void mach_ref(const Geo &g, const Dev &d, const Params &p, float TCOOL_MIN, int NR,
                                         float &vlimit, float &rolimit) {
  const int IMID = p.IMID, KMID = p.KMID;
  float GM1 = p.GA - 1.0f;
  float TRATIO = 1.0f + p.MACHLIM * p.MACHLIM * 0.5f * GM1;
  float RFGM1 = 1.0f / GM1;
  float RORATIO = pow_r(TRATIO, RFGM1);
  const int JREF = d.JLE[NR - 1];
  const size_t n = IX(IMID, JREF, KMID);
  float TREF = d.T_STATIC[n];
  float WREFSQ = d.VX[n] * d.VX[n] + d.VR[n] * d.VR[n] + d.WT[n] * d.WT[n];
  float TOREL_REF = TREF + 0.5f * WREFSQ / p.CP;
  float TLIMIT = TOREL_REF / TRATIO;
  if (TLIMIT > TCOOL_MIN) TLIMIT = TCOOL_MIN;
  vlimit = sqrtf(2 * p.CP * (TOREL_REF - TLIMIT));
  float ROSTAG_REL = d.RO[n] * pow_r(TOREL_REF / TREF, RFGM1);
  rolimit = ROSTAG_REL / RORATIO;
}

// Density limiter on the flagged (I,K) lines, mf:5115-5122.  The workgroup
// holding the line (IMID, KMID), on which every row's reference node lies,
// then forms the per-row relative Mach limits (mf:5127-5145) into d.SC for
// k_mach_apply, one row per thread, after that line's fix.
__global__ void k_rolim_fix(Geo g, Dev d, StepScalars s, Params p) {
  if (gated(d)) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < g.IM * g.KM && d.FLAGS[t] != 0) {
    const int I = t % g.IM + 1, K = t / g.IM + 1;
    d.FLAGS[t] = 0;  // ready for the next step's check
    for (int J = 3; J <= g.JM - 3; J++)
      d.RO[IX(I, J, K)] = rolim_one(d.RO[IX(I, J, K)], d.RO[IX(I, J - 2, K)], d.RO[IX(I, J - 1, K)],
                                    d.RO[IX(I, J + 1, K)], d.RO[IX(I, J + 2, K)], s.ROLIM);
  }
  const int tmid = (p.IMID - 1) + g.IM * (p.KMID - 1);
  if (tmid / (int)blockDim.x != (int)blockIdx.x) return;
  __syncthreads();
  for (int NR = threadIdx.x + 1; NR <= g.NROWS; NR += blockDim.x)
    mach_ref(g, d, p, s.TCOOL_MIN, NR, d.SC[SC_VLIMIT + NR - 1], d.SC[SC_ROLIMIT + NR - 1]);
}

__global__ void k_mach_apply(Geo g, Dev d, Params p, StepScalars st) {
  if (gated(d)) return;
  NODE_IJK
  const size_t n = IX(I, J, K);
  const int NR = d.NROW[J - 1];
  const float V_LIMIT = d.SC[SC_VLIMIT + NR - 1], RO_LIMIT = d.SC[SC_ROLIMIT + NR - 1];
  float ro = d.RO[n];
  if (ro < RO_LIMIT) {
    ro = RO_LIMIT;
    d.RO[n] = ro;
  }
  const float vx = d.VX[n], vr = d.VR[n], wt = d.WT[n];
  float WRELSQ = vx * vx + vr * vr + wt * wt;
  float WREL = sqrtf(WRELSQ);
  float FACSAFE = V_LIMIT / WREL;
  if (FACSAFE < 1.0f) {
    d.ROVX[n] = ro * vx * FACSAFE;
    d.ROVR[n] = ro * vr * FACSAFE;
    float ROVTREL = ro * wt * FACSAFE;
    d.ROWT[n] = ROVTREL;
    float rovt = ROVTREL + ro * d.UBLADE[JKX(J, K)];
    d.ROVT[n] = rovt;
    d.RORVT[n] = rovt * d.R[JKX(J, K)];
  }
}

// ---------------------------------------------------------------------------
// NEW_MIXPLAN, mf:14922-15251 (ideal gas, ITIMST<5): one thread per
// (mixing plane, K) line; the supersonic Prandtl-Meyer sweeps are serial in I.

// MIX_BCONDS (mf:11906-12080) with MASS_AVG (mf:11729-11760): one wavefront
// per (plane, quantity).  Plane 0 is J = 1, plane p >= 1 is J = JMIX(p).
// The node quantity of each cell's four corners on the plane, weighted by
// the cell's positive mass flow -FLOWX, summed over I per cell line K (FP64
// lane sums), divided by the line's flow; the line values then moved to the
// nodes by the reference's running average and end extrapolations, and the
// whole plane's average.  The node quantities are the reference's FP32
// expressions (transcendentals through FP64, rounded once).
__device__ __forceinline__ float mixb_quantity(const Geo &g, const Dev &d, const Params &p, int q, size_t n) {
  const float RADDEG = 180.f / 3.14159265f;
  const float VX = d.VX[n], VT = d.VT[n], VR = d.VR[n];
  switch (q) {
    case 0: {
      const float D = .5f * (VX * VX + VT * VT + VR * VR);
      const float TSTAG = d.HO[n] / p.CP;
      float TS = TSTAG - D / p.CP;
      if (TS < 1.f) TS = 1.f;
      return d.P[n] * pow_r(TSTAG / TS, p.GA / (p.GA - 1));
    }
    case 1: return d.P[n];
    case 2: return d.HO[n] / p.CP;
    case 3: {
      const float D = .5f * (VX * VX + VT * VT + VR * VR);
      return (d.HO[n] - D) / p.CP;
    }
    case 4: return VT;
    case 5: return sqrtf(VX * VX + VR * VR);
    case 6: return atan2_r(VT, sqrtf(VX * VX + VR * VR)) * RADDEG;
    default: return atan2_r(VR, sqrtf(VX * VX + VR * VR)) * RADDEG;
  }
}
__global__ void __launch_bounds__(64) k_mix_bconds(Geo g, Dev d, Params p) {
  if (gated(d)) return;
  const int pl = blockIdx.x / 8, q = blockIdx.x % 8, lane = threadIdx.x;
  const int J = pl == 0 ? 1 : d.JMIX[pl - 1];
  const int KM = g.KM;
  extern __shared__ double S[];  // KM + 1 line values
  double sumall = 0.0, flowall = 0.0;
  for (int K = 1; K <= KM - 1; K++) {
    double sas = 0.0, sf = 0.0;
    for (int I = lane + 1; I <= g.IMM1; I += 64) {
      float DFLOW = -d.FLOWX[IX(I, J, K)];
      if (DFLOW < 0.0f) DFLOW = 0.0f;
      const float f = mixb_quantity(g, d, p, q, IX(I, J, K)) + mixb_quantity(g, d, p, q, IX(I + 1, J, K)) +
                      mixb_quantity(g, d, p, q, IX(I + 1, J, K + 1)) + mixb_quantity(g, d, p, q, IX(I, J, K + 1));
      sas += (double)DFLOW;
      sf += (double)DFLOW * (double)f * 0.25;
    }
    sas = wave_sum_d(sas);
    sf = wave_sum_d(sf);
    if (lane == 0) {
      sumall += sf;
      flowall += sas;
      S[K] = sf / sas;
    }
  }
  if (lane != 0) return;
  for (int K = 2; K <= KM - 1; K++) S[K] = 0.5 * (S[K] + S[K - 1]);
  S[1] = 2.0 * S[2] - S[3];
  S[KM] = 2.0 * S[KM - 1] - S[KM - 2];
  double *o = d.MIXB + (size_t)blockIdx.x * (KM + 1);
  for (int K = 1; K <= KM; K++) o[K - 1] = S[K];
  o[KM] = sumall / flowall;
}

// NEWBCONDS (mf:12085-12115), the repeating-stage condition: the inlet
// stagnation pressure, yaw angle and tangential velocity relaxed by RFINBC
// towards the exit plane's profiles (MIX_BCONDS of the last plane, in the
// reference's FP32), shifted to the deck's mid-span inlet values; the
// sines and cosines of the inlet angles recomputed.  One thread per K.
__global__ void k_newbconds(Geo g, Dev d, Params p) {
  if (gated(d)) return;
  const int K = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (K > g.KM) return;
  const float DEGRAD = 3.14159265f / 180.f;
  const double *ex = d.MIXB + (size_t)g.NROWS * 8 * (g.KM + 1);  // the last plane
  auto prof = [&](int q, int k) { return (float)ex[(size_t)q * (g.KM + 1) + k - 1]; };
  const float RFINBC1 = 1.0f - p.RFINBC;
  const float po = RFINBC1 * d.PO1[K - 1] + p.RFINBC * (p.PO_IN_MID + prof(0, K) - prof(0, p.KMID));
  const float bs = RFINBC1 * d.BS[K - 1] + p.RFINBC * prof(6, K);
  const float vt = RFINBC1 * d.VTIN[K - 1] + p.RFINBC * (p.VT_IN_MID + prof(4, K) - prof(4, p.KMID));
  d.PO1[K - 1] = po;
  d.BS[K - 1] = bs;
  d.VTIN[K - 1] = vt;
  d.RPO1[K - 1] = 1.f / po;
  d.BSSIN[K - 1] = sin_r(bs * DEGRAD);
  d.BSCOS[K - 1] = cos_r(bs * DEGRAD);
  d.BRSIN[K - 1] = sin_r(d.BR[K - 1] * DEGRAD);
  d.BRCOS[K - 1] = cos_r(d.BR[K - 1] * DEGRAD);
}

// The trailing-edge separation force of rows with IF_CUSP = 2
// (mf:5184-5233), after NEW_MIXPLAN: one workgroup per (row, K).  From
// NUP_I1 points upstream of the trailing edge to N_WAKE points downstream,
// the nodes I = 1..6 that lie more than SEP_THIK x the blade thickness (at
// the later of the two start points) below the I = 1 surface extrapolated
// from its start point get their meridional momentum and their angular
// momentum relative to the blade scaled by SEP_DRAG; then the nodes
// I = IM-6..IM above the extrapolated I = IM surface, from NUP_IM points
// upstream.  The two passes run in the reference's order (a node in both is
// scaled twice).  The host has checked that every index stays in the grid.
__global__ void __launch_bounds__(256) k_te_sep(Geo g, Dev d) {
  if (gated(d)) return;
  const int NR = blockIdx.x / g.KM + 1, K = blockIdx.x % g.KM + 1;
  if (d.IF_CUSP[NR - 1] != 2) return;
  const int IM = g.IM, JTEDGE = d.JTE[NR - 1], NWKE = d.N_WAKE[NR - 1];
  const int JSEP_I1 = JTEDGE - d.NUP_I1[NR - 1], JSEP_IM = JTEDGE - d.NUP_IM[NR - 1];
  const int JTHIK = max(JSEP_I1, JSEP_IM);
  const float FDRAG = d.SEP_DRAG[NR - 1];
  const float THIK_LIM = d.SEP_THIK[NR - 1] * d.RT_THICK[JKX(JTHIK, K)];
  const float *RT = d.RTHETA, *X = d.X;
  const float SSSLOPE = (RT[IX(1, JSEP_I1 + 1, K)] - RT[IX(1, JSEP_I1 - 1, K)]) /
                        (X[JKX(JSEP_I1 + 1, K)] - X[JKX(JSEP_I1 - 1, K)]);
  const float PSSLOPE = (RT[IX(IM, JSEP_IM + 1, K)] - RT[IX(IM, JSEP_IM - 1, K)]) /
                        (X[JKX(JSEP_IM + 1, K)] - X[JKX(JSEP_IM - 1, K)]);
  auto drag = [&](size_t n, int J) {
    d.ROVX[n] = FDRAG * d.ROVX[n];
    d.ROVR[n] = FDRAG * d.ROVR[n];
    const float RO_VBLADE = d.RO[n] * d.UBLADE[JKX(J, K)];
    const float RO_WT = d.ROVT[n] - RO_VBLADE;
    const float RO_VT = FDRAG * RO_WT + RO_VBLADE;
    d.RORVT[n] = RO_VT * d.R[JKX(J, K)];
  };
  const int n1 = (JTEDGE + NWKE - JSEP_I1 + 1) * 6;
  for (int t = threadIdx.x; t < n1; t += blockDim.x) {
    const int J = JSEP_I1 + t / 6, I = t % 6 + 1;
    const float RTEXTRAP = RT[IX(1, JSEP_I1, K)] + SSSLOPE * (X[JKX(J, K)] - X[JKX(JSEP_I1, K)]);
    const size_t n = IX(I, J, K);
    if ((RTEXTRAP - RT[n]) > THIK_LIM) drag(n, J);
  }
  __syncthreads();
  const int n2 = (JTEDGE + NWKE - JSEP_IM + 1) * 7;
  for (int t = threadIdx.x; t < n2; t += blockDim.x) {
    const int J = JSEP_IM + t / 7, I = IM - 6 + t % 7;
    const float RTEXTRAP = RT[IX(IM, JSEP_IM, K)] + PSSLOPE * (X[JKX(J, K)] - X[JKX(JSEP_IM, K)]);
    const size_t n = IX(I, J, K);
    if ((RT[n] - RTEXTRAP) > THIK_LIM) drag(n, J);
  }
}

__global__ void k_mixplan(Geo g, Dev d, Params p, float *scratch) {
  if (gated(d)) return;
  extern __shared__ float ml[];  // 9 * IM floats: per-I work arrays of this line
  const int t = blockIdx.x;      // one wavefront per (mixing plane, K) line
  const int NR = t / g.KM + 1, K = t % g.KM + 1;
  if (NR > g.NROWS - 1) return;
  const int IM = g.IM, IMID = p.IMID, lane = threadIdx.x;
  float *ALPHAA = ml, *GAMA = ml + IM, *WREF = ml + 2 * IM, *ROREF = ml + 3 * IM, *PISENT = ml + 4 * IM,
        *ROISENT = ml + 5 * IM, *WMACH = ml + 6 * IM, *TISENT = ml + 7 * IM, *WISV = ml + 8 * IM;
  (void)scratch;
  const float RFMIX = p.RFMIX, RFMIX1 = 1.0f - RFMIX;
  int KSUB = K;
  if (K == g.KM) KSUB = g.KM - 1;
  int J = d.JMIX[NR - 1], JP1, JP2, JP3;
  float F_ANGLE;
  if (d.FLOWX[IX(IMID, J, KSUB)] < 0.0f) {
    JP1 = J + 1; JP2 = J + 2; JP3 = J + 3; F_ANGLE = p.FANGLE;
  } else {
    J = d.JMIX[NR - 1] + 1;
    JP1 = J - 1; JP2 = J - 2; JP3 = J - 3; F_ANGLE = 0.0f;
  }
  const float F_ANGLE1 = 1.0f - F_ANGLE;
  // mixed-out state from the mid-pitch point of the plane, mf:14968-14995
  const float RJP1 = d.R[JKX(JP1, K)], UJP1 = d.UBLADE[JKX(JP1, K)];
  const size_t nm = IX(IMID, JP1, K);
  const float ROMIX = d.RO[nm];
  const float VXMIX = d.ROVX[nm] / ROMIX;
  const float VRMIX = d.ROVR[nm] / ROMIX;
  const float VTMIX = d.RORVT[nm] / ROMIX / RJP1;
  const float ROEMIX = d.ROE[nm];
  const float VMSQ = VXMIX * VXMIX + VRMIX * VRMIX;
  const float VABSQ = VMSQ + VTMIX * VTMIX;
  const float WTMIX = VTMIX - UJP1;
  const float WMIXSQ = VMSQ + WTMIX * WTMIX;
  const float WMIX = sqrtf(WMIXSQ);
  const float ALPHAMIX = asin_r(WTMIX / WMIX);
  const float GAMAMIX = atan2_r(VRMIX, VXMIX);
  const float TMIX = (ROEMIX / ROMIX - 0.5f * VABSQ) / p.CV;
  const float ROTHALPY = p.CP * TMIX + 0.5f * (WMIXSQ - UJP1 * UJP1);
  const float GAGAS = p.GA;
  const float PMIX = ROMIX * p.RGAS * TMIX;
  const float RJP2 = d.R[JKX(JP2, K)], UNOWJP2 = d.UBLADE[JKX(JP2, K)];
  const float RJP3 = d.R[JKX(JP3, K)], UNOWJP3 = d.UBLADE[JKX(JP3, K)];
  // isentropic state and flow angles at JP2 for every I, mf:15000-15125
  int nmach = 0;
  for (int I = lane + 1; I <= IM; I += 64) {
    const size_t n2 = IX(I, JP2, K), n3 = IX(I, JP3, K);
    float ROJP2 = d.RO[n2];
    float XVELJP2 = d.ROVX[n2] / ROJP2;
    float RVELJP2 = d.ROVR[n2] / ROJP2;
    float TVELJP2 = d.RORVT[n2] / RJP2 / ROJP2;
    float WTJP2 = TVELJP2 - UNOWJP2;
    float VMSQJP2 = XVELJP2 * XVELJP2 + RVELJP2 * RVELJP2;
    float VELSQJP2 = VMSQJP2 + TVELJP2 * TVELJP2;
    float WACT_SQ_JP2 = VMSQJP2 + WTJP2 * WTJP2;
    float WACT_JP2 = sqrtf(WACT_SQ_JP2);
    float T_ACT_JP2 = (d.ROE[n2] / ROJP2 - 0.5f * VELSQJP2) / p.CV;
    float ROJP3 = d.RO[n3];
    float ROVXJP3 = d.ROVX[n3], ROVRJP3 = d.ROVR[n3];
    float ROVTJP3 = d.RORVT[n3] / RJP3;
    float ROWTJP3 = ROVTJP3 - UNOWJP3 * ROJP3;
    float ROVMSQJP3 = ROVXJP3 * ROVXJP3 + ROVRJP3 * ROVRJP3;
    float ROWRELSQJP3 = ROVMSQJP3 + ROWTJP3 * ROWTJP3;
    float ROWRELJP3 = sqrtf(ROWRELSQJP3);
    float ALPHAJP3 = asin_r(ROWTJP3 / ROWRELJP3);
    float GAMAJP3 = atan2_r(ROVRJP3, ROVXJP3);
    float PSTAT = ROJP2 * p.RGAS * T_ACT_JP2;
    float TISENT_JP2 = TMIX * pow_r(PSTAT / PMIX, p.FGA);
    float WIS_SQ_JP2 = 2 * (ROTHALPY - p.CP * TISENT_JP2 + 0.5f * UNOWJP2 * UNOWJP2);
    float ROISENT_JP2 = PSTAT / p.RGAS / TISENT_JP2;
    ROISENT[I - 1] = ROISENT_JP2;
    PISENT[I - 1] = PSTAT;
    TISENT[I - 1] = TISENT_JP2;
    if (WIS_SQ_JP2 < 1.0f) WIS_SQ_JP2 = 1.0f;
    WISV[I - 1] = sqrtf(WIS_SQ_JP2);
    WREF[I - 1] = WACT_JP2;
    ROREF[I - 1] = ROJP2;
    float wm = WACT_JP2 / sqrtf(GAGAS * p.RGAS * T_ACT_JP2);
    WMACH[I - 1] = wm;
    if (wm > 1.0f) nmach++;
    ALPHAA[I - 1] = F_ANGLE1 * ALPHAMIX + F_ANGLE * ALPHAJP3;
    GAMA[I - 1] = F_ANGLE1 * GAMAMIX + F_ANGLE * GAMAJP3;
  }
  const bool supersonic = __any(nmach != 0);
  __syncthreads();
  // Prandtl-Meyer angle variation for supersonic relative flow, serial in I,
  // mf:15139-15173, then periodicity of the angle, mf:15177-15182
  if (supersonic && lane == 0) {
    float FROTN = 1.0f;
    if (WTMIX < 0.0f) FROTN = -1.0f;
    for (int I = IMID + 1; I <= IM; I++) {
      const int a = I - 1, b = I - 2;
      if (WMACH[a] > 1.001f && WMACH[b] > 1.001f) {
        float DELP = PISENT[a] - PISENT[b];
        float DELTAI = sqrtf(WMACH[a] * WMACH[a] - 1) / (ROREF[a] * WREF[a] * WREF[a]);
        float DELTAIM1 = sqrtf(WMACH[b] * WMACH[b] - 1) / (ROREF[b] * WREF[b] * WREF[b]);
        ALPHAA[a] = ALPHAA[b] + FROTN * DELP * 0.5f * (DELTAI + DELTAIM1);
      }
    }
    for (int I = IMID - 1; I >= 1; I--) {
      const int a = I - 1, b = I;
      if (WMACH[a] > 1.001f && WMACH[b] > 1.001f) {
        float DELP = PISENT[a] - PISENT[b];
        float DELTAI = sqrtf(WMACH[a] * WMACH[a] - 1) / (ROREF[a] * WREF[a] * WREF[a]);
        float DELTAIP1 = sqrtf(WMACH[b] * WMACH[b] - 1) / (ROREF[b] * WREF[b] * WREF[b]);
        ALPHAA[a] = ALPHAA[b] + FROTN * DELP * 0.5f * (DELTAI + DELTAIP1);
      }
    }
    float ROTN = ALPHAA[IM - 1] - ALPHAA[0];
    float SUMFP = 0.0f;
    for (int I = 1; I <= IM; I++) {
      ALPHAA[I - 1] = ALPHAA[I - 1] - ROTN * (SUMFP - 0.5f);
      SUMFP = SUMFP + d.FP[I - 1];
    }
  }
  __syncthreads();
  // relax the JP2 state towards the isentropic one, mf:15199-15246
  for (int I = lane + 1; I <= IM; I += 64) {
    const size_t n2 = IX(I, JP2, K);
    float RO_JP2 = ROISENT[I - 1];
    float WT_JP2 = WISV[I - 1] * sin_r(ALPHAA[I - 1]);
    float VM_JP2 = WISV[I - 1] * cos_r(ALPHAA[I - 1]);
    float VX_JP2 = VM_JP2 * cos_r(GAMA[I - 1]);
    float VR_JP2 = VM_JP2 * sin_r(GAMA[I - 1]);
    float VT_JP2 = WT_JP2 + UNOWJP2;
    float V_JP2_SQ = VM_JP2 * VM_JP2 + VT_JP2 * VT_JP2;
    float ROE_JP2 = ROISENT[I - 1] * (p.CV * TISENT[I - 1] + 0.5f * V_JP2_SQ);
    d.RO[n2] = RFMIX1 * d.RO[n2] + RFMIX * RO_JP2;
    d.ROVX[n2] = RFMIX1 * d.ROVX[n2] + RFMIX * RO_JP2 * VX_JP2;
    d.ROVR[n2] = RFMIX1 * d.ROVR[n2] + RFMIX * RO_JP2 * VR_JP2;
    d.RORVT[n2] = RFMIX1 * d.RORVT[n2] + RFMIX * RO_JP2 * VT_JP2 * RJP2;
    d.ROE[n2] = RFMIX1 * d.ROE[n2] + RFMIX * ROE_JP2;
  }
}

// ---------------------------------------------------------------------------
// every-5-step checks: VRMS/VMAX (mf:3846-3864), CHECK_FLOW/ECONT
// (mf:5264-5283), EMAX/EAVG (mf:5291-5317)

// Sum and max of W^2 over J>=2: stage 1 per block in FP64.
__global__ void k_vrms_part(Geo g, Dev d, double *part) {
  if (gated(d)) return;
  __shared__ double ssum[256];
  __shared__ float smax[256];
  const int tid = threadIdx.x;
  const size_t N = (size_t)g.IM * g.JM * g.KM;
  double s = 0.0;
  float m = 0.0f;
  // wavefronts walk whole (J,K) lines, lanes the I points of a line
  const int lane = tid & 63, nw = blockDim.x >> 6, nlines = g.JM * g.KM;
  for (int ln = blockIdx.x * nw + wave_id(); ln < nlines; ln += gridDim.x * nw) {
    const int J = ln % g.JM + 1, K = ln / g.JM + 1;
    if (J < 2) continue;
    for (int I = lane + 1; I <= g.IM; I += 64) {
      const size_t n = IX(I, J, K);
      float EKE = d.VR[n] * d.VR[n] + d.WT[n] * d.WT[n] + d.VX[n] * d.VX[n];
      s += (double)EKE;
      if (EKE > m) m = EKE;
    }
  }
  (void)N;
  ssum[tid] = s;
  smax[tid] = m;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (tid < off) {
      ssum[tid] += ssum[tid + off];
      smax[tid] = fmaxf(smax[tid], smax[tid + off]);
    }
    __syncthreads();
  }
  if (tid == 0) {
    part[2 * blockIdx.x] = ssum[0];
    part[2 * blockIdx.x + 1] = (double)smax[0];
  }
}

__global__ void k_vrms_final(Geo g, Dev d, const double *part, int nparts, int exact) {
  if (gated(d)) return;
  __shared__ double ss[256];
  __shared__ double sm[256];
  if (exact) {
    if (threadIdx.x != 0) return;
    float VRMS = 0.0f, VMAX = 0.0f;
    for (int K = 1; K <= g.KM; K++)
      for (int J = 2; J <= g.JM; J++)
        for (int I = 1; I <= g.IM; I++) {
          const size_t n = IX(I, J, K);
          float EKE = d.VR[n] * d.VR[n] + d.WT[n] * d.WT[n] + d.VX[n] * d.VX[n];
          if (EKE > VMAX) VMAX = EKE;
          VRMS = VRMS + EKE;
        }
    d.SC[SC_VRMS] = VRMS;
    d.SC[SC_VMAX] = sqrtf(VMAX);
    return;
  }
  double s = 0.0, m = 0.0;
  for (int b = threadIdx.x; b < nparts; b += 256) {
    s += part[2 * b];
    m = fmax(m, part[2 * b + 1]);
  }
  ss[threadIdx.x] = s;
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      ss[threadIdx.x] += ss[threadIdx.x + off];
      sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + off]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    d.SC[SC_VRMS] = (float)ss[0];
    d.SC[SC_VMAX] = sqrtf((float)sm[0]);
  }
}

// per-J mass flow, reference order over I then K (exact) or FP64 (fast)
__global__ void k_checkflow(Geo g, Dev d, int exact) {
  if (gated(d)) return;
  const int J = blockIdx.x + 1;
  if (J > g.JM) return;
  float SUMAS;
  if (exact) {
    if (threadIdx.x != 0) return;
    SUMAS = 0;
    for (int I = 1; I <= g.IMM1; I++)
      for (int K = 1; K <= g.KMM1; K++) SUMAS = SUMAS - d.FLOWX[IX(I, J, K)];
  } else {
    // lanes over I, K lines in turn, FP64
    double s = 0.0;
    for (int K = 1; K <= g.KMM1; K++)
      for (int I = threadIdx.x + 1; I <= g.IMM1; I += 64) s -= (double)d.FLOWX[IX(I, J, K)];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (threadIdx.x != 0) return;
    SUMAS = (float)s;
  }
  const float NB = (float)d.NBLADE[J - 1];
  d.BLADE_FLOW[J - 1] = SUMAS * NB;
  d.CHECK_FLOW[J - 1] = (SUMAS + 0.0f) * NB + 0.0f - 0.0f;
}

// SETFLO, mf:8006-8052: a body force that pulls the mass flow of every J
// towards FLOWIN (IN_FLOW = 3) or towards the mean over J (otherwise).
// Per J, the flow through the plane scaled to the first row's pitch,
// FLOW(J) = -sum FLOWX * NBLADE(J) / NBLADE(1), summed K-major as the
// reference does (exact) or in FP64 (fast); one workgroup per J.
__global__ void __launch_bounds__(64) k_setflo_flow(Geo g, Dev d, int exact) {
  if (gated(d)) return;
  const int J = blockIdx.x + 1;
  const float NBJ = (float)d.NBLADE[J - 1], NB1 = (float)d.NBLADE[0];
  if (exact) {
    if (threadIdx.x != 0) return;
    float F = 0.0f;
    for (int K = 1; K <= g.KMM1; K++)
      for (int I = 1; I <= g.IMM1; I++) F = F - d.FLOWX[IX(I, J, K)] * NBJ / NB1;
    d.SETFLO[J - 1] = F;
    return;
  }
  double s = 0.0;
  for (int K = 1; K <= g.KMM1; K++)
    for (int I = threadIdx.x + 1; I <= g.IMM1; I += 64) s -= (double)(d.FLOWX[IX(I, J, K)] * NBJ / NB1);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (threadIdx.x == 0) d.SETFLO[J - 1] = (float)s;
}

// The target flow and, per J, the meridional mass-flux change DELTA(J) at
// the mid-passage node (mf:8022-8036), into d.SETFLO[JM + J - 1].  One
// workgroup; the mean over J is sequential (exact) or FP64 (fast).
__global__ void __launch_bounds__(256) k_setflo_delta(Geo g, Dev d, Params p, int exact) {
  if (gated(d)) return;
  __shared__ double part[256];
  __shared__ float avflow;
  const int JM = g.JM, t = threadIdx.x;
  if (p.IN_FLOW == 3) {
    if (t == 0) avflow = p.FLOWIN / (float)d.NBLADE[0];
  } else if (exact) {
    if (t == 0) {
      float A = 0.0f;
      for (int J = 1; J <= JM; J++) A = A + d.SETFLO[J - 1];
      avflow = A / (float)JM;
    }
  } else {
    double s = 0.0;
    for (int J = t + 1; J <= JM; J += 256) s += (double)d.SETFLO[J - 1];
    part[t] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
      if (t < off) part[t] += part[t + off];
      __syncthreads();
    }
    if (t == 0) avflow = (float)part[0] / (float)JM;
  }
  __syncthreads();
  for (int J = t + 1; J <= JM; J += 256) {
    const size_t n = IX(p.IMID, J, p.KMID);
    const float ROVMER = sqrtf(d.ROVX[n] * d.ROVX[n] + d.ROVR[n] * d.ROVR[n]);
    float DELTA = (avflow / d.SETFLO[J - 1] - 1.0f) * ROVMER;
    DELTA = DELTA * p.RFLOW;
    d.SETFLO[JM + J - 1] = DELTA;
  }
}

// The change along the local meridional direction of every node, with the
// relative swirl momentum scaled alike, mf:8037-8049
__global__ void k_setflo_apply(Geo g, Dev d) {
  if (gated(d)) return;
  NODE_IJK
  const unsigned n = IX(I, J, K);
  const float DELTA = d.SETFLO[g.JM + J - 1];
  const float rovx = d.ROVX[n], rovr = d.ROVR[n];
  float ROVMER = sqrtf(rovx * rovx + rovr * rovr);
  if (rovx < 0.0f) ROVMER = -ROVMER;
  d.ROVX[n] = rovx + rovx / ROVMER * DELTA;
  d.ROVR[n] = rovr + rovr / ROVMER * DELTA;
  const float RO_WT = d.ROVT[n] - d.UBLADE[JKX(J, K)] * d.RO[n];
  d.ROVT[n] = d.ROVT[n] + RO_WT / ROVMER * DELTA;
}

__global__ void k_econt(Geo g, Dev d) {
  if (gated(d)) return;
  __shared__ float sm[256];
  float e = 0.0f;
  const float cf1 = d.CHECK_FLOW[0];
  for (int J = threadIdx.x + 1; J <= g.JM; J += 256) {
    float RATIO = d.CHECK_FLOW[J - 1] / cf1;
    e = fmaxf(e, fabsf(1.f - RATIO));
  }
  sm[threadIdx.x] = e;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) sm[threadIdx.x] = fmaxf(sm[threadIdx.x], sm[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    d.SC[SC_ECONT] = sm[0];
    d.SC[SC_FLOWRAT] = d.BLADE_FLOW[0];
    d.SC[SC_FLOWOUT] = d.BLADE_FLOW[g.JM - 1];
  }
}

__device__ __forceinline__ float dvmer_at(const Geo &g, const Dev &d, int I, int J, int K, float RVEF) {
  const size_t n = IX(I, J, K);
  float XD = d.DX[JKX(J, K)], RD = d.DR[JKX(J, K)], SD = d.DS[JKX(J, K)];
  float VM_START = (XD * d.VX[n] + RD * d.VR[n]) / SD;
  float VM_END = (XD * d.ROVX[n] + RD * d.ROVR[n]) / SD / d.RO[n];
  return (VM_END - VM_START) * RVEF;
}

// stage 1: per block FP64 sum of |DVMER| and the first max-|DVMER| node
__global__ void k_emax_part(Geo g, Dev d, double *part) {
  if (gated(d)) return;
  __shared__ double ssum[256];
  __shared__ float sabs[256];
  __shared__ long long sidx[256];
  __shared__ float sval[256];
  const int tid = threadIdx.x;
  const float RIJKM = 2.f / (float)(g.IM * g.JM * g.KM);
  const float VREF = sqrtf(d.SC[SC_VRMS] * RIJKM);
  const float RVEF = 100.f / VREF;
  const size_t N = (size_t)g.IM * g.JM * g.KM;
  double s = 0.0;
  float ba = -1.0f, bv = 0.0f;
  long long bi = -1;
  // wavefronts walk whole (J,K) lines, lanes the I points of a line
  const int lane = tid & 63, nw = blockDim.x >> 6, nlines = g.JM * g.KM;
  for (int ln = blockIdx.x * nw + wave_id(); ln < nlines; ln += gridDim.x * nw) {
    const int J = ln % g.JM + 1, K = ln / g.JM + 1;
    if (J < 2) continue;
    for (int I = lane + 1; I <= g.IM; I += 64) {
      const size_t n = IX(I, J, K);
      float v = dvmer_at(g, d, I, J, K, RVEF);
      s += (double)fabsf(v);
      if (fabsf(v) > ba) { ba = fabsf(v); bv = v; bi = (long long)n; }
    }
  }
  (void)N;
  ssum[tid] = s; sabs[tid] = ba; sval[tid] = bv; sidx[tid] = bi;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (tid < off) {
      ssum[tid] += ssum[tid + off];
      const float a2 = sabs[tid + off];
      const long long i2 = sidx[tid + off];
      if (a2 > sabs[tid] || (a2 == sabs[tid] && i2 >= 0 && (sidx[tid] < 0 || i2 < sidx[tid]))) {
        sabs[tid] = a2; sval[tid] = sval[tid + off]; sidx[tid] = i2;
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    part[4 * blockIdx.x] = ssum[0];
    part[4 * blockIdx.x + 1] = (double)sabs[0];
    part[4 * blockIdx.x + 2] = (double)sval[0];
    part[4 * blockIdx.x + 3] = (double)sidx[0];
  }
}

__global__ void k_emax_final(Geo g, Dev d, const double *part, int nparts, int exact) {
  if (gated(d)) return;
  __shared__ double ssum[256], sabs[256], sidx[256];
  __shared__ float sval[256];
  const float RIJKM = 2.f / (float)(g.IM * g.JM * g.KM);
  const float VREF = sqrtf(d.SC[SC_VRMS] * RIJKM);
  const float RVEF = 100.f / VREF;
  if (exact) {
    if (threadIdx.x != 0) return;
    float EMAX = 0.0f, EAVG = 0.0f;
    int IMAX = 0, JMAX = 0, KMAX = 0;
    for (int K = 1; K <= g.KM; K++)
      for (int J = 2; J <= g.JM; J++)
        for (int I = 1; I <= g.IM; I++) {
          float DVMER = dvmer_at(g, d, I, J, K, RVEF);
          EAVG = EAVG + fabsf(DVMER);
          if (fabsf(DVMER) > fabsf(EMAX)) {
            EMAX = DVMER;
            IMAX = I; JMAX = J; KMAX = K;
          }
        }
    d.SC[SC_EAVG] = EAVG / (float)(g.IM * g.JMM1 * g.KM);
    d.SC[SC_EMAX] = EMAX;
    d.SC[SC_VREF] = VREF;
    d.CTL[CTL_IMAX] = IMAX; d.CTL[CTL_JMAX] = JMAX; d.CTL[CTL_KMAX] = KMAX;
    return;
  }
  double s = 0.0, ba = -1.0, bi = -1.0;
  float bv = 0.0f;
  for (int b = threadIdx.x; b < nparts; b += 256) {
    s += part[4 * b];
    const double a = part[4 * b + 1], i = part[4 * b + 3];
    if (i >= 0 && (a > ba || (a == ba && i < bi))) { ba = a; bi = i; bv = (float)part[4 * b + 2]; }
  }
  ssum[threadIdx.x] = s; sabs[threadIdx.x] = ba; sidx[threadIdx.x] = bi; sval[threadIdx.x] = bv;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    const int t = threadIdx.x;
    if (t < off) {
      ssum[t] += ssum[t + off];
      const double a = sabs[t + off], i = sidx[t + off];
      if (i >= 0 && (a > sabs[t] || (a == sabs[t] && (sidx[t] < 0 || i < sidx[t])))) {
        sabs[t] = a; sidx[t] = i; sval[t] = sval[t + off];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    d.SC[SC_EAVG] = (float)ssum[0] / (float)(g.IM * g.JMM1 * g.KM);
    d.SC[SC_EMAX] = sval[0];
    d.SC[SC_VREF] = VREF;
    // the first node in the reference's loop order (K, J, I) with the
    // largest |DVMER|: the smallest linear index among equal values; none
    // when every value is zero (the reference keeps its last location then)
    if (sidx[0] >= 0 && sabs[0] > 0.0) {
      const long long n = (long long)sidx[0];
      d.CTL[CTL_IMAX] = (int)(n % g.IM) + 1;
      d.CTL[CTL_JMAX] = (int)((n / g.IM) % g.JM) + 1;
      d.CTL[CTL_KMAX] = (int)(n / ((long long)g.IM * g.JM)) + 1;
    }
  }
}

// ---------------------------------------------------------------------------
// STEPUP, mf:7916-8002 (ideal gas, ITIMST<5)

__global__ void k_stepup_temp(Geo g, Dev d, Params p) {
  if (gated(d)) return;
  NODE_IJK
  const size_t n = IX(I, J, K);
  float TLIM = 0.1f * d.TO1[p.KMID - 1];
  float VLIM = 0.01f * d.VM1[p.KMID - 1] * d.VM1[p.KMID - 1];
  const float vx = d.VX[n], vt = d.VT[n], vr = d.VR[n];
  float VSQ = vx * vx + vt * vt + vr * vr;
  float HSTAT = d.HO[n] - 0.5f * VSQ;
  float TSTATIC = HSTAT * p.RCP;
  float GAMNOW = p.GA;
  if (TSTATIC < TLIM) TSTATIC = TLIM;
  float WTREL = d.WT[n];
  float WSQ = VSQ - (vt * vt - WTREL * WTREL);
  if (WSQ < VLIM) WSQ = VLIM;
  float V_SONIC = sqrtf(GAMNOW * p.RGAS * TSTATIC);
  float VPLUSC = sqrtf(WSQ) + V_SONIC;
  d.TEMP1[n] = p.VSOUND / VPLUSC;
}

__global__ void k_stepup_step(Geo g, Dev d, float RELAX) {
  if (gated(d)) return;
  NODE_IJK
  if (I > g.IMM1 || K > g.KMM1 || J < 2) return;
  const int IP1 = I + 1, KP1 = K + 1;
  const float RELAX1 = 1.0f - RELAX;
  const float *T = d.TEMP1;
  float AMACHP = T[IX(I, J - 1, K)] + T[IX(I, J - 1, KP1)] + T[IX(IP1, J - 1, K)] + T[IX(IP1, J - 1, KP1)];
  if (J == 2) AMACHP = T[IX(I, 1, K)] + T[IX(IP1, 1, K)] + T[IX(I, 1, KP1)] + T[IX(IP1, 1, KP1)];
  float AMACHL = T[IX(I, J, K)] + T[IX(I, J, KP1)] + T[IX(IP1, J, K)] + T[IX(IP1, J, KP1)];
  float AVMACH = 0.125f * (AMACHL + AMACHP);
  const size_t n = IX(I, J, K);
  float STEPNEW = d.BSTEP[n] * AVMACH;
  float st = RELAX * STEPNEW + RELAX1 * d.STEP[n];
  d.STEP[n] = st;
  d.RSTEP[n] = st / d.BSTEP[n];
}

__global__ void k_stepup_mix(Geo g, Dev d) {
  if (gated(d)) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int per = g.IMM1 * g.KMM1;
  const int NR = t / per + 1, c = t % per;
  if (NR > g.NROWS - 1) return;
  const int I = c % g.IMM1 + 1, K = c / g.IMM1 + 1;
  const int J = d.JMIX[NR - 1];
  d.STEP[IX(I, J + 1, K)] = d.STEP[IX(I, J + 2, K)];
  d.RSTEP[IX(I, J + 1, K)] = 0.0f;
}

// ---------------------------------------------------------------------------
// mass-flux-weighted station integrals (EFICOOL mf:10685-10739 + SUMFLX
// mf:7883-7910), FP64 per-J sums: one workgroup per J

__global__ void k_integrals(Geo g, Dev d, Params p) {
  if (gated(d)) return;
  const int J = blockIdx.x + 1;
  if (J > g.JM) return;
  __shared__ double sh[7][256];
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  const float TO1M = d.TO1[p.KMID - 1], PO1M = d.PO1[p.KMID - 1];
  const float NB = (float)d.NBLADE[J - 1];
  auto temps = [&](int I, int K, float &tstag, float &prat0, float &rvt, float &ent, float &ts, float &ps) {
    const size_t n = IX(I, J, K);
    float EKE = 0.5f * (d.VX[n] * d.VX[n] + d.VT[n] * d.VT[n] + d.VR[n] * d.VR[n]);
    float TS = (d.HO[n] - EKE) * p.RCP;
    float TSTAG = d.HO[n] * p.RCP;
    if (TSTAG < 0.1f * TO1M) TSTAG = 0.1f * TO1M;
    if (TS < 0.1f * TO1M) TS = 0.1f * TO1M;
    float PSTAT = d.P[n];
    if (PSTAT < 0.02f * PO1M) PSTAT = 0.02f * PO1M;
    float PSTAG = PSTAT * pow_r(TSTAG / TS, 1.0f / p.FGA);
    float TRAT = TS / TO1M, PRAT = PSTAT / PO1M, ENTRPY = 0.0f;
    if (!((PRAT > 0.99999f && PRAT < 1.00001f) || (TRAT > 0.99999f && TRAT < 1.00001f)))
      ENTRPY = p.CP * log_r(TRAT) - p.RGAS * log_r(PRAT);
    tstag = TSTAG; prat0 = PSTAG / PO1M; rvt = d.RORVT[n] / d.RO[n]; ent = ENTRPY; ts = TS; ps = d.P[n];
  };
  for (int c = threadIdx.x; c < g.IMM1 * g.KMM1; c += blockDim.x) {
    const int I = c % g.IMM1 + 1, K = c / g.IMM1 + 1;
    double DFLOW = -(double)d.FLOWX[IX(I, J, K)] * (double)NB;
    double s[6] = {0, 0, 0, 0, 0, 0};
    const int ci[4] = {I, I + 1, I + 1, I}, ck[4] = {K, K, K + 1, K + 1};
    for (int q = 0; q < 4; q++) {
      float a, b, c2, e, f, h;
      temps(ci[q], ck[q], a, b, c2, e, f, h);
      s[0] += b; s[1] += a; s[2] += c2; s[3] += e; s[4] += f; s[5] += h;
    }
    acc[0] += DFLOW;
    for (int q = 0; q < 6; q++) acc[q + 1] += 0.25 * DFLOW * s[q];
  }
  for (int q = 0; q < 7; q++) sh[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off)
      for (int q = 0; q < 7; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 7; q++) d.INTEG[(size_t)q * g.JM + J - 1] = sh[q][0];
}

// EFICOOL's stage and machine figures, uncooled branch (mf:11253-11548,
// IFGAS = 0), in FP64 from the SUMFLX integrals in d.INTEG (k_integrals:
// BLADE_FLOW, SUMPO, SUMTO, SUMRVT, SUM_ENTPY, SUMTSTAT, SUMPSTAT per J) and,
// for the pressure reactions, FP-weighted pitchwise sums of P at K = 1,
// KM/2 and KM (mf:11299-11357).  One wavefront.  Record r has the planes
// tab[4r .. 4r+2] = J1, J2M1, JMIX and tab[4r+3] = its stage (0 = the whole
// machine); it writes kEffFields doubles at out[r * kEffFields].
__global__ void __launch_bounds__(64) k_eficool(Geo g, Dev d, Params p, const int *tab, int nrec, double *out) {
  const int lane = threadIdx.x, JM = g.JM;
  const double *FLOW = d.INTEG, *SPO = FLOW + JM, *STO = FLOW + 2 * JM, *STS = FLOW + 5 * JM, *SPS = FLOW + 6 * JM;
  const double CP = p.CP, GA = p.GA, FGA = (GA - 1.0) / GA, PO1M = d.PO1[p.KMID - 1];
  const int KMIDE = g.KM / 2 > 0 ? g.KM / 2 : 1;
  const int ks[3] = {1, KMIDE, g.KM};
  for (int r = 0; r < nrec; r++) {
    const int J1 = tab[4 * r], J2M1 = tab[4 * r + 1], JMIX = tab[4 * r + 2], stage = tab[4 * r + 3];
    // pitchwise static pressures: [plane J1, JMIX, J2M1][K = 1, KM/2, KM]
    double pp[3][3];
    const int js[3] = {J1, JMIX, J2M1};
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        double s = 0.0;
        for (int I = lane + 1; I <= g.IMM1; I += 64)
          s += 0.5 * ((double)d.P[IX(I, js[a], ks[b])] + (double)d.P[IX(I + 1, js[a], ks[b])]) * (double)d.FP[I - 1];
        pp[a][b] = wave_sum_d(s);
      }
    if (lane != 0) continue;
    const double POIN = SPO[J1 - 1] / FLOW[J1 - 1] * PO1M, PO2 = SPO[J2M1 - 1] / FLOW[J2M1 - 1] * PO1M;
    const double PSIN = SPS[J1 - 1] / FLOW[J1 - 1], PSOUT = SPS[J2M1 - 1] / FLOW[J2M1 - 1];
    const double TO_IN = STO[J1 - 1] / FLOW[J1 - 1], TO_OUT = STO[J2M1 - 1] / FLOW[J2M1 - 1];
    const double TORAT = TO_OUT / TO_IN, DELTO = TO_OUT - TO_IN;
    double PORAT = PO2 / POIN;
    const double PSRAT = PSOUT / POIN;
    if (PORAT < 0.01) PORAT = 0.01;
    const double DELTOIS = pow(PORAT, FGA) * TO_IN - TO_IN, DELT_TOT_STAT = pow(PSRAT, FGA) * TO_IN - TO_IN;
    double ETAIS = 1.0, ETATS = 1.0;
    if (DELTO > 0.0) { ETAIS = DELTOIS / DELTO; ETATS = DELT_TOT_STAT / DELTO; }
    if (DELTO < 0.0) { ETAIS = DELTO / DELTOIS; ETATS = DELTO / DELT_TOT_STAT; }
    double EPOLY = log(TORAT) / log(PORAT) / FGA;
    if (TORAT > 1.0) EPOLY = 1.0 / EPOLY;
    double REACTION = NAN, RH = NAN, RM = NAN, RT = NAN, WORK;
    if (stage != 0) {
      const double tin = STS[J1 - 1] / FLOW[J1 - 1], tmid = STS[JMIX - 1] / FLOW[JMIX - 1],
                   tout = STS[J2M1 - 1] / FLOW[J2M1 - 1];
      REACTION = (tmid - tout) / (tin - tout);
      if (PSOUT > PSIN) REACTION = 1.0 - REACTION;
      RH = (pp[1][0] - pp[2][0]) / (pp[0][0] - pp[2][0]);
      RM = (pp[1][1] - pp[2][1]) / (pp[0][1] - pp[2][1]);
      RT = (pp[1][2] - pp[2][2]) / (pp[0][2] - pp[2][2]);
      if (pp[2][1] > pp[0][1]) { RH = 1.0 - RH; RM = 1.0 - RM; RT = 1.0 - RT; }
      WORK = CP * (FLOW[J1 - 1] * TO_IN - FLOW[J2M1 - 1] * TO_OUT) / 1000.0;
    } else {
      WORK = CP * (FLOW[0] * TO_IN - FLOW[JM - 2] * TO_OUT) / 1000.0;
    }
    double *o = out + (size_t)r * kEffFields;
    const double v[kEffFields] = {POIN, PO2, PSIN, PSOUT, TO_IN, TO_OUT, FLOW[J1 - 1], FLOW[J2M1 - 1], PORAT,
                                  ETAIS, ETATS, EPOLY, REACTION, RH, RM, RT, WORK, DELTO >= 0.0 ? 1.0 : 0.0};
    for (int f = 0; f < kEffFields; f++) o[f] = v[f];
  }
}

}  // namespace mh

namespace mh {

// ---------------------------------------------------------------------------
// SET_XLENGTH (mf:12194-12517): wall distances and mixing-length scales, once
// per run.  k_xl_dist: one thread per node.  The end walls: the nearest hub
// (casing) grid point over J +/- 20, then the distance along that wall
// point's face normal (mf:12227-12287); the blade surfaces I = 1 and I = IM
// of the node's own row: the nearest surface point over J +/- 20 (clipped to
// the row) and K +/- 5 in true 3D distance, then the normal distance there,
// or the point distance off the blade (mf:12296-12402); the two combined as
// the reference does (mf:12408-12424).  FP32 in the reference's statement
// order; sinf is the device's.
__global__ void __launch_bounds__(256) k_xl_dist(Geo g, Dev d, const float *TH, float *DIST, int IBOUND) {
  if (gated(d)) return;
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.IM * g.JM * g.KM) return;
  const int I = n % g.IM + 1, J = (n / g.IM) % g.JM + 1, K = n / (g.IM * g.JM) + 1;
  const int NR = d.NROW[J - 1];
  const float XP = d.X[JKX(J, K)], RP = d.R[JKX(J, K)];
  const float PITCH = 6.283185f * RP / (float)d.NBLADE[J - 1];
  const float DREF = 0.5f * PITCH;
  const int IW = g.IMID, j1 = max(J - 20, 2), j2 = min(J + 20, g.JM);
  auto endwall = [&](int KW) {
    float dmin = 1.0e10f;
    int jmin = j1;
    for (int jw = j1; jw <= j2; jw++) {
      const float xd = XP - d.X[JKX(jw, KW)], rd = RP - d.R[JKX(jw, KW)];
      const float ds = xd * xd + rd * rd;
      if (ds < dmin) {
        dmin = ds;
        jmin = jw;
      }
    }
    const float ax = d.ASX[IX(IW, jmin, KW)], ar = d.ASR[IX(IW, jmin, KW)];
    const float atot = sqrtf(ax * ax + ar * ar);
    const float xn = ax / atot, rn = ar / atot;
    const float xd = XP - d.X[JKX(jmin, KW)], rd = RP - d.R[JKX(jmin, KW)];
    return fabsf(xd * xn + rd * rn);
  };
  float hub = endwall(1);
  if (IBOUND == 1 || IBOUND > 2) hub = DREF;
  float tip = endwall(g.KM);
  if (IBOUND >= 2) tip = DREF;
  const float EW = hub * tip / (hub + tip);
  const int jst = d.JSTART[NR - 1], jle = d.JLE[NR - 1], jte = d.JTE[NR - 1], jmx = d.JMIX[NR - 1];
  int b1 = J - 20;
  if (b1 < 2) b1 = 2;
  if (b1 < jst) b1 = jst;
  if (b1 >= jte) b1 = jte - 1;
  int b2 = J + 20;
  if (b2 > g.JM) b2 = g.JM;
  if (b2 > jmx) b2 = jmx;
  if (b2 <= jle) b2 = jle + 1;
  const int k1 = max(K - 5, 1), k2 = min(K + 5, g.KM);
  const float TP = TH[n];
  float sd[2];
  for (int side = 0; side < 2; side++) {
    const int IS = side ? g.IM : 1;
    float dmin = 1.0e10f;
    int jm = 0, km = 0;
    for (int js = b1; js <= b2; js++)
      for (int ks = k1; ks <= k2; ks++) {
        const float RS = d.R[JKX(js, ks)];
        const float xd = XP - d.X[JKX(js, ks)], rd = RP - RS;
        const float s = sinf((TP - TH[IX(IS, js, ks)]) / 2.0f);
        const float td = 4.0f * RP * RS * (s * s);
        const float ds = xd * xd + rd * rd + td;
        if (ds < dmin) {
          dmin = ds;
          jm = js;
          km = ks;
        }
      }
    if (jm > 0) {
      const int kn = km == g.KM ? g.KM - 1 : km;
      const float bx = d.ABX[IX(IS, jm, kn)], br = d.ABR[IX(IS, jm, kn)], bt = d.ABT[JKX(jm, kn)];
      const float atot = sqrtf(bx * bx + br * br + bt * bt);
      const float xn = bx / atot, rn = br / atot, tn = bt / atot;
      const float RS = d.R[JKX(jm, km)];
      const float xd = XP - d.X[JKX(jm, km)], rd = RP - RS;
      const float tdif = 2.0f * sqrtf(RP * RS) * sinf((TP - TH[IX(IS, jm, km)]) / 2.0f);
      sd[side] = fabsf(xd * xn + rd * rn + tdif * tn);
      if (jm <= jle || jm >= jte) sd[side] = sqrtf(xd * xd + rd * rd + tdif * tdif);
    } else {
      sd[side] = DREF;
    }
  }
  const float BD = sd[0] * sd[1] / (sd[0] + sd[1]);
  float dm = 0.0f;
  if (K > 1 && K < g.KM) {
    float q = sqrtf(EW * EW + BD * BD);
    if (q < 1.0e-10f) q = 1.0e-10f;
    dm = BD * EW / q;
  }
  if (g.KM == 2) dm = BD;
  if (g.IM == 2) dm = EW;
  DIST[n] = dm;
  (void)I;
}

// The mixing-length limit along each row at (IMID, KMID), varying linearly in
// meridional distance from inlet to leading edge, leading to trailing edge,
// trailing edge to mixing plane (mf:12428-12449).  XL4 = [IN | LE | TE | DN]
// per row.  One thread per J.
__global__ void k_xl_limit(Geo g, Dev d, const float *DIST, const float *XL4, float *XLIMIT) {
  if (gated(d)) return;
  const int J = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (J > g.JM) return;
  const int NR = d.NROW[J - 1], K = g.KMID, nr = g.NROWS;
  const float xin = XL4[NR - 1], xle = XL4[nr + NR - 1], xte = XL4[2 * nr + NR - 1], xdn = XL4[3 * nr + NR - 1];
  const int jst = d.JSTART[NR - 1], jle = d.JLE[NR - 1], jte = d.JTE[NR - 1], jmx = d.JMIX[NR - 1];
  const float *S = d.SMERID;
  float xllim = 0.0f;
  if (J <= jle) xllim = xin + (S[JKX(J, K)] - S[JKX(jst, K)]) / (S[JKX(jle, K)] - S[JKX(jst, K)]) * (xle - xin);
  if (J > jle && J <= jte)
    xllim = xle + (S[JKX(J, K)] - S[JKX(jle, K)]) / (S[JKX(jte, K)] - S[JKX(jle, K)]) * (xte - xle);
  if (J > jte) xllim = xte + (S[JKX(J, K)] - S[JKX(jte, K)]) / (S[JKX(jmx, K)] - S[JKX(jte, K)]) * (xdn - xte);
  XLIMIT[J - 1] = 2.0f * DIST[IX(g.IMID, J, K)] * xllim;
}

// The blended mixing length, the free-stream damping ratio and their cell
// averages (mf:12460-12508): XLENGTH = (0.41 l)^2, FST_RAT, DWALLSQ per cell,
// zero on the planes that hold no cell and on each row's first J.  One thread
// per node.
__global__ void __launch_bounds__(256) k_xl_cells(Geo g, Dev d, const float *DIST, const float *XLIMIT, float dblend,
                                                  float expon, float fblend) {
  if (gated(d)) return;
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.IM * g.JM * g.KM) return;
  const int I = n % g.IM + 1, J = (n / g.IM) % g.JM + 1, K = n / (g.IM * g.JM) + 1;
  if (I > g.IMM1 || J < 2 || K > g.KMM1) {
    d.XLENGTH[n] = 0.0f;
    d.FST_RAT[n] = 0.0f;
    d.DWALLSQ[n] = 0.0f;
    return;
  }
  float xl[8], t3[8], dd[8];
  const int di[8] = {0, 0, 0, 0, 1, 1, 1, 1}, dj[8] = {0, 0, -1, -1, 0, 0, -1, -1}, dk[8] = {0, 1, 0, 1, 0, 1, 0, 1};
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int Jq = J + dj[q];
    const float lim = XLIMIT[Jq - 1];
    float x = DIST[IX(I + di[q], Jq, K + dk[q])];
    dd[q] = x;
    const float ratio = x / lim;
    if (ratio > dblend) x = lim * (1.0f - fblend * powf(ratio, -expon));
    xl[q] = x;
    const float r = x / lim;
    t3[q] = ((r * r) * r) * r;  // the reference compiler's order for r**4
  }
  const float axl = 0.125f * (xl[0] + xl[1] + xl[2] + xl[3] + xl[4] + xl[5] + xl[6] + xl[7]);
  const float aft = 0.125f * (t3[0] + t3[1] + t3[2] + t3[3] + t3[4] + t3[5] + t3[6] + t3[7]);
  const float ads = 0.125f * (dd[0] + dd[1] + dd[2] + dd[3] + dd[4] + dd[5] + dd[6] + dd[7]);
  const int NR = d.NROW[J - 1];
  const bool first = J == d.JSTART[NR - 1];
  d.XLENGTH[n] = first ? 0.0f : 0.41f * 0.41f * axl * axl;
  d.FST_RAT[n] = first ? 0.0f : aft * d.FSTURB[NR - 1];
  d.DWALLSQ[n] = ads * ads;
}

}  // namespace mh
