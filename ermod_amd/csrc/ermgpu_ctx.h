// ermgpu_ctx.h -- internal to libermgpu.so: the context, device buffers,
// error and launch helpers, and the launchers of the register-FFT kernels
// (compiled per precision and grid length in ermgpu_pme.hip).
#pragma once

#include "ermgpu.h"

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ermgpu_kern.h"

namespace ermh {
using namespace erm;

struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

// stage events per batch (start, then the end of each stage) and the
// stage times erm_stage_times reports (nine stages and the total)
constexpr int NSTAGE_EV = 10, NSTAGE_OUT = 10;

}  // namespace ermh

// ============================================================== the context
struct erm_ctx {
  using DBuf = ermh::DBuf;
  using Params = erm::Params;
  erm_params P{};
  std::string err;
  hipStream_t st = nullptr;
  hipEvent_t ev[8] = {};
  double stage_ms[ermh::NSTAGE_OUT] = {};
  bool fp32 = false;
  int Bmax = 256;
  // topology
  bool have_topo = false, have_bins = false, have_frame = false;
  int natom = 0, nmol = 0, ntype = 0;
  std::vector<int> mol_begin, ljtype, sluvid, uvspec;
  std::vector<double> charge, lj_eps, lj_sig2, mol_charge, sitemass, bfcoord;
  int slvmax = 0;
  std::vector<int> tagpt, kof, sltlist;
  int slt_template = -1, ns = 0;
  int nvt = 0;
  std::vector<int> vtypes;  // compact index -> ljtype
  double total_charge = 0.0;
  // bins
  int numslv = 0, ermax = 0, esmax = 0;
  std::vector<int> uvmax, uvsoft, specoff;
  std::vector<double> uvcrd, escrd;
  // frame
  double cell[9] = {}, invcl[9] = {}, volume = 0.0, L[3] = {};
  bool have_prevcl = false;
  double prevcl[9] = {};
  double stat_weight_system = 1.0;
  std::vector<int> jof_k;  // sorted index of tagpt index k
  bool jof_stale = true;   // jof_k not yet read back for this frame
  Params kp{};
  // refs rigid: realcal_self of the first configuration is reused
  bool have_usreal = false;
  // device buffers
  DBuf d_mol_begin, d_charge, d_tagpt, d_cvt, d_mol_charge, d_uvspec, d_sitemass, d_bf, d_qslt, d_qslt_t, d_sltlj,
      d_sltlj64;
  DBuf d_pos, d_cellid, d_count, d_cell_start, d_fill, d_rm, d_mol_k, d_mol_ns, d_mol_first, d_mol_q,
      d_mol_spec, d_mol_m, d_mol_d1, d_jof, d_kof;
  DBuf d_sx, d_sy, d_sz, d_sq, d_sct, d_sg, d_sw;
  DBuf d_zcount, d_zstart, d_zfill, d_zfine, d_zsite, d_zpos, d_zg, d_zq, d_zw, d_esite;
  int nsites = 0;             // sites of the physical molecules
  int slab_S = 0, slab_R = 0; // fused PME slab: planes per workgroup, registers per thread
  DBuf d_tw1, d_tw2, d_tw3, d_engfac;
  DBuf d_edges, d_specoff, d_uvmax, d_uvsoft, d_escrd;
  DBuf d_w3h;
  // solute gradient: per-frame solvent potential grid and batch buffers
  DBuf d_ggrid, d_gHk, d_gtab, d_gq_t, d_gqd, d_gepsd, d_gpart, d_gphi, d_gub, d_gw, d_gout;
  uint64_t frame_serial = 0, ggrid_serial = ~0ULL;
  // batching: device-side call arguments, per-batch stage events, graph replay
  DBuf d_args;
  uint64_t *h_args = nullptr;      // pinned {seed, first_index}
  bool args_ready = false;         // d_args holds the current call's arguments
  std::vector<hipEvent_t> bev;     // 8 events per batch of the last call
  hipEvent_t *cur_ev = nullptr;    // events of the batch being enqueued
  int wsol_ones = 0;               // d_wsol holds 1.0 for this many configurations
  bool capturing = false, use_graph = false;
  hipGraphExec_t gexec = nullptr;
  std::string gkey;
  uint64_t alloc_gen = 0;          // bumped by every device (re)allocation
  int ggrid_excl = -2;
  int xslab_Q = 0;            // planes per x-c2r group of the register path
  int tune_S = 16, tune_Q = 8;  // caps on slab depth and planes per group (autotuned)
  int xslab_T = erm::NTHR;           // threads per x-slab workgroup, 256 or 512 (autotuned)
  int zcol_T = erm::NTHR;            // threads per z-column workgroup, 256 or 512 (autotuned)
  bool xslab_yf = false;        // inverse y transform inside the x slab (autotuned)
  int pair_path = 0;            // 0 tiled pair kernel, 1 one thread per candidate molecule
  int ncu = 256;                // compute units of the device (persistent grids)
  double slt_radius = 8.0;      // test particle's largest site distance from its centroid
  double mol_radius = 2.0;      // solvent molecules' largest site distance from the first site
  std::vector<int> zst;         // first z-sorted site of every plane (host copy)
  size_t lds_cap = 150 * 1024; // dynamic LDS a kernel may request (160 KB per CU)
  // insertion modes: host molecules of the aggregate centre, and the centre
  std::vector<int> hostmol;
  DBuf d_hostmol, d_aggcen;
  // trajectory frames: FP32 staging and the permutation
  DBuf d_tstage, d_tperm;
  // intermediate state: empty reference bins, accumulators, weights
  DBuf d_zerodns, d_edintm, d_uvintm, d_intm_sc, d_wi;
  // reference structure (RefInfo, insertion.F90:623-723): host atoms and
  // their reference coordinates and weights, the solute's reference
  // coordinates and weights; per frame the fitted and best-fit solute
  DBuf d_flexbf;  // body configurations of a flexible test particle
  bool have_ref = false, gauss_init = false;
  std::vector<int> ref_atoms;
  std::vector<double> ref_host, ref_hw, ref_slt, ref_sw;
  DBuf d_refatoms, d_refgather, d_refsite, d_refw;
  DBuf d_spos, d_sref, d_sg3, d_sw3, d_sqv, d_w1h, d_w2h, d_Hk, d_phi, d_selfp, d_uv, d_uvout, d_hist,
      d_wsol, d_w, d_flceng, d_sltj, d_selfreal, d_usreal;
  DBuf d_edens, d_ecorr, d_eself, d_slnuv, d_mnmx, d_scal, d_status, d_cellm;
  std::vector<DBuf *> all() {
    return {&d_mol_begin, &d_charge, &d_tagpt, &d_cvt, &d_mol_charge, &d_uvspec, &d_sitemass, &d_bf, &d_qslt, &d_qslt_t,
            &d_sltlj, &d_sltlj64, &d_pos, &d_cellid, &d_count, &d_cell_start, &d_fill, &d_rm, &d_mol_k, &d_mol_ns,
            &d_mol_first, &d_mol_q, &d_mol_spec, &d_mol_m, &d_mol_d1, &d_jof, &d_kof, &d_sx, &d_sy, &d_sz, &d_sq, &d_sct, &d_sg,
            &d_sw, &d_zcount, &d_zstart, &d_zfill, &d_zfine, &d_zsite, &d_zpos, &d_zg, &d_zq, &d_zw, &d_esite, &d_tw1, &d_tw2, &d_tw3, &d_engfac, &d_edges, &d_specoff, &d_uvmax, &d_uvsoft, &d_escrd,
            &d_w3h, &d_args, &d_ggrid, &d_gHk, &d_gtab, &d_gq_t, &d_gqd, &d_gepsd, &d_gpart, &d_gphi, &d_gub, &d_gw,
            &d_gout, &d_spos, &d_sref, &d_sg3, &d_sw3, &d_sqv, &d_w1h, &d_w2h, &d_Hk, &d_phi, &d_selfp, &d_uv,
            &d_uvout, &d_hist, &d_wsol, &d_w, &d_flceng, &d_sltj, &d_selfreal, &d_usreal, &d_edens, &d_ecorr,
            &d_eself, &d_slnuv, &d_mnmx, &d_scal, &d_status, &d_cellm, &d_hostmol, &d_aggcen, &d_tstage,
            &d_tperm, &d_zerodns, &d_edintm, &d_uvintm, &d_intm_sc, &d_wi, &d_refatoms, &d_refgather,
            &d_refsite, &d_refw, &d_flexbf};
  }
};

namespace ermh {

inline int fail(erm_ctx *c, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIPCHK(c, call)                                                                        \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess) return fail((c), ERM_E_DEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

inline int ensure(erm_ctx *c, DBuf &b, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (b.bytes >= bytes) return ERM_OK;
  b.release();
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    return fail(c, ERM_E_DEVICE, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  }
  b.bytes = bytes;
  ++c->alloc_gen;
  return ERM_OK;
}

template <typename X> inline int upload(erm_ctx *c, DBuf &b, const X *src, size_t n) {
  int s = ensure(c, b, n * sizeof(X));
  if (s) return s;
  if (n) HIPCHK(c, hipMemcpyAsync(b.p, src, n * sizeof(X), hipMemcpyHostToDevice, c->st));
  return ERM_OK;
}

template <typename X> inline X *dp(DBuf &b) { return reinterpret_cast<X *>(b.p); }

inline int sync(erm_ctx *c) {
  HIPCHK(c, hipStreamSynchronize(c->st));
  return ERM_OK;
}

inline int launch_check(erm_ctx *c, const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(c, ERM_E_DEVICE, "launch %s: %s", what, hipGetErrorString(e));
  return ERM_OK;
}

// raises the dynamic LDS limit of kernel f to `bytes` when it exceeds the
// default 64 KB; a refused request is cleared from the runtime's error state
inline int set_lds(erm_ctx *c, const void *f, size_t bytes, const char *what) {
  if (bytes <= 64 * 1024) return ERM_OK;
  hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(c, ERM_E_UNSUPPORTED, "%s: %zu bytes of LDS refused (%s)", what, bytes, hipGetErrorString(e));
  }
  return ERM_OK;
}

#define TRY(x)              \
  do {                      \
    int s_ = (x);           \
    if (s_) return s_;      \
  } while (0)

inline const char *status_name(int code) {
  switch (code) {
    case ERM_E_ENG_MIN: return "eng_min: energy at or below the lowest bin edge";
    case ERM_E_ENG_SFT: return "eng_sft: energy above the soft region";
    case ERM_E_MBIN: return "mbin: energy in the overflow bin below 1e3";
    case ERM_E_ENG_CNS: return "eng_cns: solvent molecule without species";
    case ERM_E_INS_LIMIT: return "no reference-structure insertion accepted in 4096 attempts";
    default: return "error";
  }
}

// first error raised by the device during the call (configuration index
// within the call), from the status words at d_status (the context's own
// by default); synchronises the stream
inline int check_status(erm_ctx *c, const int *d_status = nullptr) {
  int h[8];
  HIPCHK(c, hipMemcpyAsync(h, d_status ? d_status : c->d_status.p, sizeof h, hipMemcpyDeviceToHost, c->st));
  TRY(sync(c));
  if (h[0]) {
    double v;
    memcpy(&v, &h[4], sizeof v);
    return fail(c, h[0], "%s (configuration %d, sorted molecule %d, energy %.6g)", status_name(h[0]), h[1],
                h[2], v);
  }
  return ERM_OK;
}

// stage event i of the batch being enqueued (none while a graph is captured)
inline int mark(erm_ctx *c, int i) {
  if (c->capturing || !c->cur_ev) return ERM_OK;
  HIPCHK(c, hipEventRecord(c->cur_ev[i], c->st));
  return ERM_OK;
}

// whether the x slab takes the inverse y transform (512-thread slabs, K2 = K1)
inline bool xslab_folds_y(const erm_ctx *c) { return c->xslab_yf && c->xslab_T == 512 && c->kp.K2 == c->kp.K1; }

// Launchers of the register-FFT kernels, one instantiation per precision T
// and grid length K (ermgpu_pme.hip): z columns (spread + z FFT + Green's
// function, or the spectral z sum), inverse y columns, the x slab with the
// gather, and the solvent-potential transforms of the adjoint.
template <typename T, int K> int launch_zcol_reg(erm_ctx *c, const Params &k, int n, int nblk);
// (launch_zcol_fft from site site0 and configuration cfg0 of the batch
// buffers, k.ns sites per configuration; launch_zcol_fft_runs over a run
// table with a persistent grid)
template <typename T, int K>
int launch_zcol_fft(erm_ctx *c, const Params &k, int n, int nblk, size_t site0 = 0, size_t cfg0 = 0);
template <typename T, int K>
int launch_zcol_fft_runs(erm_ctx *c, const Params &k, int nblk, const GRun *runs, int nrun, int nitem, int *counter);
template <typename T, int K> void launch_ycol_reg(erm_ctx *c, const Params &k, int n);
template <typename T, int K> int launch_xslab(erm_ctx *c, const Params &k, int n, size_t lds);
template <typename T, int K> void launch_xr2c(erm_ctx *c, const Params &k);
template <typename T, int K> void launch_xc2r(erm_ctx *c, const Params &k);
template <typename T, int K, int SIGN> void launch_ydense(erm_ctx *c, const Params &k);
template <typename T, int K> void launch_zgreen(erm_ctx *c, const Params &k);

}  // namespace ermh
