// Batched, grouped gfx950 kernels for the spaQR per-front hot path.
//
// Every kernel takes a device array of task descriptors and runs one 256-thread workgroup
// (four 64-wide wavefronts) per descriptor, so all fronts / blocks of one tree level, of
// uneven sizes, go into one launch.  Storage is column-major; every T factor is kept in
// 32-column blocks (T is NB x k, one NB x NB upper-triangular block per panel), the layout
// LAPACK's dgeqrt/dtpqrt use with NB = 32.
#pragma once
#include <cstdint>
#include <hip/hip_runtime.h>

namespace spaqr {

constexpr int NB = 32;    // reflector block (panel) width
constexpr int WG = 256;   // threads per workgroup

// geqrt of A (m x n, lda), T (NB x n) out.  Also used for the interface scaling QR.
template <typename S> struct GeqrtTask { S* A; S* T; int m, n, lda, pad; };

// C (m x n, ldc) <- Q^T C, Q from a geqrt (V m x k in Vd with ldv, T NB x k).
// Optionally followed by C <- C R^{-1} (R upper k2 x k2 in Rd, ldr) when Rd != nullptr.
template <typename S> struct LarfbTask {
    const S* V; const S* T; S* C; const S* R;
    int m, n, k, ldv, ldc, k2, ldr, pad;
};

// tpqrt chain of one interior: R (k x k upper, ldr) against each B_j (rows_j x k).
template <typename S> struct TpqrtLink { S* B; S* T; int rows, ldb; };
template <typename S> struct TpqrtTask { S* R; int k, ldr, link0, nlinks; };

// Cooperative tpqrt chain: workgroups b0 .. b0 + ceil(k/32) - 1 of the launch, workgroup p
// owning column panel p of R and of every link's B; cnt points at the chain's ceil(k/32)
// zero-initialised progress counters (links finished per panel).
template <typename S> struct TpqrtCoopTask { S* R; int* cnt; int k, ldr, link0, nlinks, b0, pad; };

// tpmqrt chain of one (interior c, column block m, column tile starting at col0):
// [Atop (k x ncols); B_j(:, col0:col0+ncols)] <- Q_j^T [...] for each link j in order
// (Atop already points at the tile).
template <typename S> struct TpmqrtLink { const S* V; const S* T; S* B; int rows, ldv, ldb, pad; };
template <typename S> struct TpmqrtTask { S* Atop; int k, lda, ncols, link0, nlinks, col0; };

// geqp3 (column-pivoted QR with early stop at the sparsification tolerance) of the
// gathered C (k x w, ld k).  Outputs rank, V (k x rank in Vout, ld k), tau, T (NB x rank).
template <typename S> struct Geqp3Task {
    S* C; S* Vout; S* tau; S* T; S* vn1; S* vn2; int* jpvt; int* rank_out;
    int k, w, pad0, pad1;
};

// Cooperative geqp3: one interface's C (k x w, ld k, in HBM) is split into column slices
// of `cpb` columns held in the LDS of `P` workgroups (blocks b0 .. b0+P-1 of a cooperative
// launch).  xch is per-task exchange scratch in HBM (geqp3_coop_xch_elems values), bar a
// zero-initialised barrier counter.  Outputs as Geqp3Task (jpvt_out: first `rank` pivots).
template <typename S> struct Geqp3CoopTask {
    const S* C; S* Vout; S* tau; S* T; int* rank_out; int* jpvt_out; S* xch; int* bar;
    int k, w, b0, P, cpb, kmax;
};
inline size_t geqp3_coop_xch_elems(int k, int P) { return 2 * (size_t)P * (k + 4); }

// One panel step of a large front, split in two launches: the panel QR (one workgroup per
// front) and the trailing update over column chunks (many workgroups per front).
//   PENT = false (geqrt): the panel is A (m x nbp, lda) from its diagonal entry, T out (ld NB).
//   PENT = true (tpqrt link): the panel is [R (nbp x nbp, ldr = lda); B (m x nbp, ldb)].
template <typename S> struct PanelTask { S* A; S* B; S* T; int m, nbp, lda, ldb; };
// [Atop (kb rows, lda); C (m rows, ldc)] <- H^T [...] over ncols columns (Atop unused when
// PENT = false); V (m x kb, ldv) and T (kb x kb, ld NB) are the panel's reflectors.
template <typename S> struct PanelApplyTask { const S* V; const S* T; S* Atop; S* C; int m, kb, ldv, lda, ldc, ncols; };
// One step of a host-driven blocked QR: the previous panel applied to some columns, then
// (qr != 0) the panel in those columns factored in the same workgroup.
template <typename S> struct PanelStepTask { PanelApplyTask<S> a; PanelTask<S> q; int apply, qr; };

// Gather/scatter of a sub-block: dst(i, j) = src(i, j) or src(j, i) (transpose).
template <typename S> struct CopyTask { const S* src; S* dst; int rows, cols, lds, ldd, trans, pad; };

// Sparsify apply of one interface-interface block A (mc x nc, lda):
// A <- A Q_n (if Vn), A[0:mc, 0:rn] <- Q_c^T A[0:mc, 0:rn] (if Vc),
// then dst(rpos + i, cpos + j) = A(i, j) for i < rc, j < rn (ldd), or the identity if ident.
// pad selects the phases: 0 all, 1 the right apply only, 2 the left apply and the scatter.
template <typename S> struct SparsApplyTask {
    S* A; const S* Vn; const S* Tn; const S* Vc; const S* Tc; S* dst;
    int mc, nc, lda, kn, rn, kc, rc, ldvn, ldvc, ldd, ident, pad;
};

// ---- preconditioner apply (vectors)
// x (m) <- Q^T x (trans = 1) or Q x (trans = 0), Q from V (m x k, ldv) with blocked T.
template <typename S> struct VecQTask { const S* V; const S* T; S* x; int m, k, ldv, trans; };
// [xc (k); xn_j (rows_j)] <- Q_j^T [...] along the chain of one interior.
template <typename S> struct VecChainLink { const S* V; const S* T; S* xn; int rows, ldv; };
template <typename S> struct VecChainTask { S* xc; int k, link0, nlinks, rmax; };  // rmax: the longest link's rows
// Sparsified interface forward/backward: Qs (m x k, ldq) + Ts, Qsp (k x r, ld k) + Tsp.
template <typename S> struct VecSparsTask { const S* Qs; const S* Ts; const S* Qsp; const S* Tsp; S* x; int k, ldq, r, pad; };
// x[0:k] -= sum_j A_j[0:k, 0:s_j] y_j, then x[0:k] <- R^{-1} x[0:k].
template <typename S> struct VecGemvLink { const S* A; const S* y; int lda, s; };
template <typename S> struct VecTrsvTask { const S* R; S* x; int k, ldr, link0, nlinks; };
template <typename S> struct VecCopyTask { const S* src; S* dst; int n, pad; };

// ---- least squares (rectangular systems)
// out[i] = ||A(i, 0:cols)|| (FP64) for i < rows; A points at the first row, ld lda.
template <typename S> struct RowNormTask { const S* A; double* out; int lda, rows, cols, pad; };
// One gathered row: p at its first entry (stride ld), or nullptr for a zero row.
template <typename S> struct RowSrc { const S* p; int ld, pad; };
// dst ((rows_old + nextra) x cols, ldd) = [old (rows_old x cols, lds); rows src[src0 .. src0+nextra)].
template <typename S> struct RowGatherTask { S* dst; const S* old; int ldd, lds, rows_old, cols, src0, nextra; };
// x[0:cols] -= sum_l A_l(0:rows_l, 0:cols)^T y_l (ld lda_l): the transposed off-diagonal products.
template <typename S> struct VecGemvtLink { const S* A; const S* y; int lda, rows; };
template <typename S> struct VecGemvtTask { S* x; int cols, link0, nlinks, pad; };

}  // namespace spaqr
