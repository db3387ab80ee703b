// Host-side launchers for the batched kernels (defined in kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>

#include "kernels.hpp"

namespace spaqr {

template <typename S> void launch_geqrt(const GeqrtTask<S>* d, int n, hipStream_t s);
// Blocks of at most 64 x 64: one 64-thread workgroup each; mr = register length class
// 16 / 32 / 64 >= the block's rows.
template <typename S> void launch_geqrt_wave(const GeqrtTask<S>* d, int n, int mr, hipStream_t s);
template <typename S> void launch_larfb(const LarfbTask<S>* d, int n, hipStream_t s);
// FP32 Q^T C with C resident in MFMA accumulators: m <= 32 mt (mt = 4, 8, 16), n <= 128, no trsm.
void launch_larfb_res(const LarfbTask<float>* d, int n, int mt, hipStream_t s);
// FP32 resident applies for 512 < rows <= 1024 (two waves per 32-column slab): larfb with n <= 64,
// tpmqrt chains with tiles of <= 64 columns.
void launch_larfb_res2(const LarfbTask<float>* d, int n, hipStream_t s);
void launch_tpmqrt_res2(const TpmqrtTask<float>* d, const TpmqrtLink<float>* l, int n, hipStream_t s);
// FP32 C R^{-1} with C^T resident in MFMA accumulators: k2 <= 32 mt (mt = 4, 8, 16), m <= 128, no Q^T.
void launch_trsm_res(const LarfbTask<float>* d, int n, int mt, hipStream_t s);
// FP32 sparsify apply phases with the block chunk resident in MFMA accumulators (right: pad 1,
// kn <= 32 mt, mc <= 128; left: pad 2 with Vc, kc <= 32 mt, rn <= 128).
void launch_spars_res(const SparsApplyTask<float>* d, int n, int mt, bool right, hipStream_t s);
// FP32 tpmqrt chains with each link's tile resident in MFMA accumulators: ncols <= 64, link rows <= 32 mt.
void launch_tpmqrt_res(const TpmqrtTask<float>* d, const TpmqrtLink<float>* l, int n, int mt, hipStream_t s);
template <typename S> void launch_panel_qr(const PanelTask<S>* d, int n, bool pent, hipStream_t s);
template <typename S> void launch_panel_apply(const PanelApplyTask<S>* d, int n, bool pent, hipStream_t s);
template <typename S> void launch_panel_step(const PanelStepTask<S>* d, int n, hipStream_t s);
template <typename S> void launch_tpqrt(const TpqrtTask<S>* d, const TpqrtLink<S>* l, int n, hipStream_t s);
// Chains with k <= 64 and links of at most 64 rows: one 64-thread workgroup per chain.
template <typename S> void launch_tpqrt_wave(const TpqrtTask<S>* d, const TpqrtLink<S>* l, int n, hipStream_t s);
// Cooperative chains (one workgroup per 32-column panel; blk2task maps each of the nblocks
// workgroups to its chain).  A launch must not exceed tpqrt_coop_max_blocks() workgroups.
template <typename S>
void launch_tpqrt_coop(const TpqrtCoopTask<S>* d, const int* blk2task, int nblocks, const TpqrtLink<S>* l, int* err,
                       hipStream_t s);
template <typename S> int tpqrt_coop_max_blocks();
template <typename S> void launch_tpmqrt(const TpmqrtTask<S>* d, const TpmqrtLink<S>* l, int n, hipStream_t s);
// Tasks with k <= 64, tiles of <= 64 columns and links of 1..64 rows (mr = register length
// class 16 / 32 / 64 >= the longest link): one 64-thread workgroup each.
template <typename S>
void launch_tpmqrt_wave(const TpmqrtTask<S>* d, const TpmqrtLink<S>* l, int n, int mr, hipStream_t s);
// parts: workgroups per task (large blocks spread over several)
template <typename S> void launch_copy(const CopyTask<S>* d, int n, hipStream_t s, int parts = 1);
template <typename S> void launch_geqp3(const Geqp3Task<S>* d, int n, double tol, double tie_rel, hipStream_t s);
template <typename S> void launch_spars_apply(const SparsApplyTask<S>* d, int n, hipStream_t s);
// Blocks of at most 64 x 64 (mr = register length class 16 / 32 / 64 >= the reflector
// dimension): one wave per task.  Larfb tasks are Q^T C column chunks (trsm = false) or
// C R^{-1} row chunks (trsm = true).
template <typename S> void launch_larfb_small(const LarfbTask<S>* d, int n, int mr, bool trsm, hipStream_t s);
template <typename S> void launch_spars_small(const SparsApplyTask<S>* d, int n, int mr, hipStream_t s);
template <typename S> void launch_fill(S* p, size_t n, S v, hipStream_t s);
template <typename S> void launch_vec_q(const VecQTask<S>* d, int n, hipStream_t s);
template <typename S> void launch_vec_chain(const VecChainTask<S>* d, const VecChainLink<S>* l, int n, hipStream_t s);
// Chains with k <= 64 whose links all have 1..64 rows: one wave per chain.
// Chains with k <= 512 whose links have 1..256 * nrw rows (nrw = 1 or 2): one workgroup per
// chain, panels pipelined.
template <typename S>
void launch_vec_chain_pipe(const VecChainTask<S>* d, const VecChainLink<S>* l, int n, int nrw, hipStream_t s);
template <typename S> void launch_vec_chain_wave(const VecChainTask<S>* d, const VecChainLink<S>* l, int n, hipStream_t s);
template <typename S> void launch_vec_spars(const VecSparsTask<S>* d, int n, int fwd, hipStream_t s);
template <typename S> void launch_vec_trsv(const VecTrsvTask<S>* d, const VecGemvLink<S>* l, int n, hipStream_t s);
template <typename S> void launch_vec_copy(const VecCopyTask<S>* d, int n, hipStream_t s);
template <typename S> void launch_convert_d2s(int n, const double* x, S* y, hipStream_t s);
template <typename S> void launch_convert_s2d(int n, const S* x, double* y, hipStream_t s);

// dynamic LDS bytes per cooperative geqp3 workgroup such that blocks_per_cu of them fit on a CU
template <typename S>
size_t geqp3_coop_lds_budget(int blocks_per_cu);
template <typename S> int geqp3_coop_max_blocks(size_t lds_bytes, bool hbm);
// hbm: column slices stay in the gathered C in HBM (for interfaces too large for the LDS layout)
template <typename S>
hipError_t launch_geqp3_coop(const Geqp3CoopTask<S>* d_tasks, const int* d_blk2task, int* d_err, int nblocks,
                             double tol, double tie_rel, size_t lds_bytes, bool hbm, hipStream_t s,
                             bool cooperative = true);

// least squares (rectangular systems)
template <typename S> void launch_row_norms(const RowNormTask<S>* d, int n, hipStream_t s);
template <typename S> void launch_row_gather(const RowGatherTask<S>* d, const RowSrc<S>* src, int n, hipStream_t s);
// dst[di[q]] = si[q] < 0 ? 0 : src[si[q]], q < n
template <typename S>
void launch_vec_gather(const S* src, S* dst, const long long* si, const long long* di, int n, hipStream_t s);
// x <- R^{-T} x (the links of the task are not used)
template <typename S> void launch_vec_trsv_t(const VecTrsvTask<S>* d, int n, hipStream_t s);
template <typename S> void launch_vec_gemvt(const VecGemvtTask<S>* d, const VecGemvtLink<S>* l, int n, hipStream_t s);
// x[0:k] <- Qsp^T R_s^{-T} x[0:k]
template <typename S> void launch_vec_spars_t(const VecSparsTask<S>* d, int n, hipStream_t s);

void launch_spmv(int n, const int* rowptr, const int* col, const double* val, const double* x, double* y,
                 hipStream_t s);
void launch_dot(int n, const double* a, const double* b, double* partial, int nblocks, double* out, int sqrt_mode,
                hipStream_t s);
void launch_axpy_dev(int n, const double* h, const double* x, double* y, hipStream_t s);
void launch_scal_dev(int n, const double* h, const double* x, double* y, hipStream_t s);
void launch_axpby(int n, double a, const double* x, double b, double* y, hipStream_t s);

// Batches of systems (krylov.hip): system s owns positions [off[s], off[s + 1]); off is a device
// array of nseg + 1 ints.  partial needs seg_partials(nseg) doubles.
int seg_partials(int nseg);
// out[s * ostride] = <a, b> on system s (its square root if sqrt_mode)
// One system, fused CGS2 (krylov.hip): v <- v - Q Q^T v twice over Q(:, 0:nvec) (columns of
// length n); hcol(0:nvec) receives the sum of both passes' coefficients, normpart (mdot_blocks(n)
// values) the chunk sums of the final v's squares; part holds nvec * mdot_blocks(n) partials,
// c nvec coefficients.  Then the device Givens step (column k of H, ld ldh, rotations cs / sn,
// right-hand side beta, est[0] = |beta[k + 1]| / ref_norm), the triangular solve for y and
// x <- x + V(:, 0:kk) y.
int mdot_blocks(int n);
void launch_cgs2(int n, const double* Q, int nvec, double* v, double* part, double* c, double* hcol, double* normpart,
                 hipStream_t s);
void launch_givens_step(int k, int nb, const double* normpart, double* hcol, double* H, int ldh, double* cs, double* sn,
                        double* beta, double ref_norm, double* est, hipStream_t s);
void launch_lsq_update(int n, int kk, const double* H, int ldh, const double* beta, double* y, const double* V,
                       double* x, hipStream_t s);
void launch_seg_dot(int nseg, const int* off, const double* a, const double* b, double* partial, double* out,
                    int ostride, int sqrt_mode, hipStream_t s);
// y <- y - h[s * hstride] x
void launch_seg_axpy(int nseg, const int* off, const double* h, int hstride, const double* x, double* y, hipStream_t s);
// y <- x / h[s * hstride] where keep[s] and h > 0, else y <- 0
void launch_seg_scal(int nseg, const int* off, const double* h, int hstride, const int* keep, const double* x,
                     double* y, hipStream_t s);
// x <- x + coef[s * cstride] v
void launch_seg_axpy_coef(int nseg, const int* off, const double* coef, int cstride, const double* v, double* x,
                          hipStream_t s);

}  // namespace spaqr
