// ermgpu_pme.hip -- the register-FFT kernels of one precision and grid
// length: z columns, inverse y columns, the x slab with the gather, and the
// adjoint's solvent-potential transforms.  Compiled once per (ERM_T, ERM_K)
// so that the instantiations build in parallel; ermgpu.hip dispatches on the
// grid length at run time.
#include <algorithm>

#include "ermgpu_ctx.h"
#include "ermgpu_grad.h"

#ifndef ERM_T
#error "ERM_T (float or double) and ERM_K (grid length) select the instantiation"
#endif

namespace ermh {

template <typename T, int K, int XT>
int launch_zcol_reg_t(erm_ctx *c, const Params &k, int n, int nblk) {
  using V = typename C2<T>::type;
  const size_t lds = zcol_reg_lds<T>(k.H, k.K2, K, XT);
  TRY(set_lds(c, (const void *)k_pme_zcol_reg<T, K, XT>, lds, "k_pme_zcol_reg"));
  hipLaunchKernelGGL((k_pme_zcol_reg<T, K, XT>), dim3(nblk, n), dim3(XT), lds, c->st, k, dp<T>(c->d_engfac),
                     dp<T>(c->d_sqv), dp<V>(c->d_w1h), dp<V>(c->d_w2h), dp<V>(c->d_w3h), dp<V>(c->d_tw3),
                     dp<V>(c->d_Hk), dp<double>(c->d_selfp));
  return ERM_OK;
}
// z columns with the tuned workgroup size (256 or 512)
template <typename T, int K> int launch_zcol_reg(erm_ctx *c, const Params &k, int n, int nblk) {
  if (c->zcol_T == 512) return launch_zcol_reg_t<T, K, 512>(c, k, n, nblk);
  return launch_zcol_reg_t<T, K, NTHR>(c, k, n, nblk);
}
template <typename T, int K, int PP, int XT>
int launch_zcol_fft_t(erm_ctx *c, const Params &k, int n, int nblk, size_t site0, size_t cfg0) {
  using V = typename C2<T>::type;
  const size_t lds = zcol_fft_lds<T>(k.H, k.K2, XT, zfft_lanes(sizeof(T), K));
  const size_t hk = cfg0 * k.K3 * (size_t)k.K2 * k.H;
  TRY(set_lds(c, (const void *)k_pme_zcol_fft<T, K, PP, XT>, lds, "k_pme_zcol_fft"));
  hipLaunchKernelGGL((k_pme_zcol_fft<T, K, PP, XT>), dim3(nblk, n), dim3(XT), lds, c->st, k, dp<T>(c->d_engfac),
                     dp<T>(c->d_sqv) + site0, dp<V>(c->d_w1h) + site0 * k.H, dp<V>(c->d_w2h) + site0 * k.K2,
                     dp<int>(c->d_sg3) + site0, dp<T>(c->d_sw3) + site0 * k.p, dp<V>(c->d_Hk) + hk,
                     dp<double>(c->d_selfp) + cfg0 * nblk);
  return ERM_OK;
}
// z columns spread and transformed along z (spline orders 4 and 6)
template <typename T, int K>
int launch_zcol_fft(erm_ctx *c, const Params &k, int n, int nblk, size_t site0, size_t cfg0) {
  if (k.p == 4) {
    if (c->zcol_T == 512) return launch_zcol_fft_t<T, K, 4, 512>(c, k, n, nblk, site0, cfg0);
    return launch_zcol_fft_t<T, K, 4, NTHR>(c, k, n, nblk, site0, cfg0);
  }
  if (c->zcol_T == 512) return launch_zcol_fft_t<T, K, 6, 512>(c, k, n, nblk, site0, cfg0);
  return launch_zcol_fft_t<T, K, 6, NTHR>(c, k, n, nblk, site0, cfg0);
}
// z columns of a heterogeneous batch (run table in device memory): a
// persistent grid (as many workgroups as fit the device at once) pulling
// items through the counter
template <typename T, int K, int PP, int XT>
int launch_zcol_runs_t(erm_ctx *c, const Params &k, int nblk, const GRun *runs, int nrun, int nitem, int *counter) {
  using V = typename C2<T>::type;
  const size_t lds = zcol_fft_lds<T>(k.H, k.K2, XT, zfft_lanes(sizeof(T), K));
  const void *f = (const void *)k_pme_zcol_fft_runs<T, K, PP, XT>;
  TRY(set_lds(c, f, lds, "k_pme_zcol_fft_runs"));
  int occ = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, f, XT, lds) != hipSuccess || occ < 1) {
    (void)hipGetLastError();
    occ = 1;
  }
  const int grid = std::min(nitem, std::max(1, occ * c->ncu));
  hipLaunchKernelGGL((k_pme_zcol_fft_runs<T, K, PP, XT>), dim3(grid), dim3(XT), lds, c->st, k, dp<T>(c->d_engfac),
                     dp<T>(c->d_sqv), dp<V>(c->d_w1h), dp<V>(c->d_w2h), dp<int>(c->d_sg3), dp<T>(c->d_sw3),
                     dp<V>(c->d_Hk), dp<double>(c->d_selfp), runs, nrun, nblk, nitem, counter);
  return ERM_OK;
}
template <typename T, int K>
int launch_zcol_fft_runs(erm_ctx *c, const Params &k, int nblk, const GRun *runs, int nrun, int nitem, int *counter) {
  if (k.p == 4) {
    if (c->zcol_T == 512) return launch_zcol_runs_t<T, K, 4, 512>(c, k, nblk, runs, nrun, nitem, counter);
    return launch_zcol_runs_t<T, K, 4, NTHR>(c, k, nblk, runs, nrun, nitem, counter);
  }
  if (c->zcol_T == 512) return launch_zcol_runs_t<T, K, 6, 512>(c, k, nblk, runs, nrun, nitem, counter);
  return launch_zcol_runs_t<T, K, 6, NTHR>(c, k, nblk, runs, nrun, nitem, counter);
}
// inverse y columns: two lanes per column in FP64 (one lane's column does
// not leave room for a second wave), one in FP32
template <typename T, int K> void launch_ycol_reg(erm_ctx *c, const Params &k, int n) {
  using V = typename C2<T>::type;
  if constexpr (sizeof(T) == 8) {
    int nblk = (2 * k.K3 * k.H + NTHR - 1) / NTHR;
    hipLaunchKernelGGL((k_pme_ycol_pair<T, K>), dim3(nblk, n), dim3(NTHR), 0, c->st, k, dp<V>(c->d_Hk));
  } else {
    int nblk = (k.K3 * k.H + NTHR - 1) / NTHR;
    hipLaunchKernelGGL((k_pme_ycol_reg<T, K>), dim3(nblk, n), dim3(NTHR), 0, c->st, k, dp<V>(c->d_tw2),
                       dp<V>(c->d_Hk));
  }
}
template <typename T, int K, int R, int XT = NTHR, bool YF = false>
int launch_xslab_r(erm_ctx *c, const Params &k, int n, size_t lds) {
  using V = typename C2<T>::type;
  TRY(set_lds(c, (const void *)k_pme_xslab<T, K, R, XT, YF>, lds, "k_pme_xslab"));
  const int nslab = (k.K3 + c->slab_S - 1) / c->slab_S;
  hipLaunchKernelGGL((k_pme_xslab<T, K, R, XT, YF>), dim3(nslab, n), dim3(XT), lds, c->st, k, c->slab_S, c->xslab_Q,
                     c->nsites, dp<V>(c->d_Hk), dp<V>(c->d_tw1), dp<int>(c->d_zstart), dp<int>(c->d_zg),
                     dp<T>(c->d_zq), dp<T>(c->d_zw), dp<int>(c->d_zsite), dp<double>(c->d_esite));
  return ERM_OK;
}
// slab_R counts sites per thread of a 256-thread workgroup; a 512-thread
// workgroup owns the same sites with half as many per thread

template <typename T, int K> int launch_xslab(erm_ctx *c, const Params &k, int n, size_t lds) {
  if (xslab_folds_y(c)) {
    if (c->slab_R == 8) return launch_xslab_r<T, K, 4, 512, true>(c, k, n, lds);
    if (c->slab_R == 16) return launch_xslab_r<T, K, 8, 512, true>(c, k, n, lds);
    return launch_xslab_r<T, K, 16, 512, true>(c, k, n, lds);
  }
  if (c->xslab_T == 512) {
    if (c->slab_R == 8) return launch_xslab_r<T, K, 4, 512>(c, k, n, lds);
    if (c->slab_R == 16) return launch_xslab_r<T, K, 8, 512>(c, k, n, lds);
    return launch_xslab_r<T, K, 16, 512>(c, k, n, lds);
  }
  if (c->slab_R == 8) return launch_xslab_r<T, K, 8>(c, k, n, lds);
  if (c->slab_R == 16) return launch_xslab_r<T, K, 16>(c, k, n, lds);
  return launch_xslab_r<T, K, 32>(c, k, n, lds);
}

template <typename T, int K> void launch_xr2c(erm_ctx *c, const Params &k) {
  using V = typename C2<T>::type;
  const int nt = k.K3 * ((k.K2 + 1) / 2);
  hipLaunchKernelGGL((k_xrow_r2c<T, K>), dim3((nt + NTHR - 1) / NTHR), dim3(NTHR), 0, c->st, k,
                     dp<T>(c->d_ggrid), dp<V>(c->d_tw1), dp<V>(c->d_gHk));
}
template <typename T, int K> void launch_xc2r(erm_ctx *c, const Params &k) {
  using V = typename C2<T>::type;
  const int nt = k.K3 * ((k.K2 + 1) / 2);
  hipLaunchKernelGGL((k_xrow_c2r<T, K>), dim3((nt + NTHR - 1) / NTHR), dim3(NTHR), 0, c->st, k,
                     dp<V>(c->d_gHk), dp<V>(c->d_tw1), dp<T>(c->d_ggrid));
}
template <typename T, int K, int SIGN> void launch_ydense(erm_ctx *c, const Params &k) {
  using V = typename C2<T>::type;
  const int nt = k.K3 * k.H;
  hipLaunchKernelGGL((k_pme_ycol_reg<T, K, SIGN>), dim3((nt + NTHR - 1) / NTHR, 1), dim3(NTHR), 0, c->st, k,
                     dp<V>(c->d_tw2), dp<V>(c->d_gHk));
}
template <typename T, int K> void launch_zgreen(erm_ctx *c, const Params &k) {
  using V = typename C2<T>::type;
  const int nt = k.H * k.K2;
  hipLaunchKernelGGL((k_zcol_green<T, K>), dim3((nt + NTHR - 1) / NTHR), dim3(NTHR), 0, c->st, k,
                     dp<T>(c->d_engfac), dp<V>(c->d_tw3), dp<V>(c->d_gHk));
}


template int launch_zcol_reg<ERM_T, ERM_K>(erm_ctx *, const Params &, int, int);
template int launch_zcol_fft<ERM_T, ERM_K>(erm_ctx *, const Params &, int, int, size_t, size_t);
template int launch_zcol_fft_runs<ERM_T, ERM_K>(erm_ctx *, const Params &, int, const GRun *, int, int, int *);
template void launch_ycol_reg<ERM_T, ERM_K>(erm_ctx *, const Params &, int);
template int launch_xslab<ERM_T, ERM_K>(erm_ctx *, const Params &, int, size_t);
template void launch_xr2c<ERM_T, ERM_K>(erm_ctx *, const Params &);
template void launch_xc2r<ERM_T, ERM_K>(erm_ctx *, const Params &);
template void launch_ydense<ERM_T, ERM_K, -1>(erm_ctx *, const Params &);
template void launch_ydense<ERM_T, ERM_K, +1>(erm_ctx *, const Params &);
template void launch_zgreen<ERM_T, ERM_K>(erm_ctx *, const Params &);

}  // namespace ermh
