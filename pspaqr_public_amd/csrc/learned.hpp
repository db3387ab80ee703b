// Randomized interface sparsification with a learned sketch size.
//
// The reference's RRQR (src/ptree.cpp:373-439) runs a full column-pivoted QR of the gathered
// C = [A(n,c)^T | A(c,m)_top] (k x w).  Here, for FP32 fronts, the rank-revealing QR runs on a
// Gaussian sketch Y = C Omega (k x s) instead, with s taken from a small MLP's rank prediction
// (the closest reference analogue is the unused adaptive range finder random_rrqr,
// src/util.cpp:643-715).  Every result is judged by the criterion geqp3 itself stops on:
// with Q from the sketch, the chosen rank r is the smallest for which every column of
// (Q^T C)[r:k, :] has norm <= tol * max_j ||C(:, j)||; an interface that fails falls back to
// geqp3 on C.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "kernels.hpp"

namespace spaqr {

constexpr int kRankFeatures = 32;
constexpr int kMlpHidden = 64;
// W1 (64 x 32, row-major), b1 (64), W2 (64 x 64), b2 (64), w3 (64), b3 (1)
constexpr int kMlpOffW1 = 0;
constexpr int kMlpOffB1 = kMlpOffW1 + kMlpHidden * kRankFeatures;
constexpr int kMlpOffW2 = kMlpOffB1 + kMlpHidden;
constexpr int kMlpOffB2 = kMlpOffW2 + kMlpHidden * kMlpHidden;
constexpr int kMlpOffW3 = kMlpOffB2 + kMlpHidden;
constexpr int kMlpOffB3 = kMlpOffW3 + kMlpHidden;
constexpr int kMlpParams = kMlpOffB3 + 1;  // 6337
constexpr int kSketchFeatureCols = 16;

// Column norms: cn[j] = ||C(:, j)||^2 (w values); *cmax2 = max(*cmax2, max_j cn[j]) through an
// atomic on the bit pattern (non-negative), so one interface may be split over column chunks
// and *cmax2 must start at zero.
template <typename S> struct ColNormTask { const S* C; S* cn; S* cmax2; int k, w; };
// Sketch Y (k x s, ld k) = C (k x w, ld k) * Omega(0:w, 0:s); one 64 x 64 tile per workgroup.
template <typename S> struct SketchTask { const S* C; S* Y; int k, w, s, pad; };
struct SketchTile { int task, r0, c0, pad; };
// Features of one interface (32 floats) from the QR'd 16-column sketch Y16 (its R diagonal),
// the column norms and the tree context.
template <typename S> struct FeatTask {
    const S* Y16; const S* cn; const S* cmax2; float* x;
    int k, w, nsout, nsin, level, pad;
};
// B = Q^T C (k x w) with Q from the sketch's first rY = *ry reflectors: *rank <- max(*rank, the
// smallest r <= rY whose tail column norms over these w columns are all <= tol * max column
// norm of the original C), INT_MAX if some column fails; an interface may be split over
// column chunks (atomic max) and *rank must start at zero.
template <typename S> struct TailTask { const S* B; const S* cmax2; const int* ry; int* rank; int k, w; };

// The tail test from the leading rows only (FP32 fronts): B (r x w, ld r) holds the first
// r = *ry rows of Q^T C and cn the squared column norms of the original C, so that
// ||B(i:, j)||^2 = cn[j] - ||B(:i, j)||^2 (Q orthogonal); otherwise as TailTask.
template <typename S> struct TailLeadTask { const S* B; const S* cn; const S* cmax2; const int* ry; int* rank; int r, w; };
// B (r x w, ld r) = M(0:r, 0:k) C (k x w): M with ld ldm, C with ld k; 64 x 64 output tiles
// (SketchTile: task, r0, c0) on FP32 MFMA.
struct TailGemmTask { const float* M; const float* C; float* B; int r, k, w, ldm; };
// M (k x k, ld ld) <- the identity, one workgroup per task.
template <typename S> struct EyeTask { S* M; int k, ld; };

template <typename S> void launch_gauss(S* om, int rows, int cols, int ld, uint64_t seed, hipStream_t s);
template <typename S> void launch_colnorm(const ColNormTask<S>* d, int n, hipStream_t s);
template <typename S>
void launch_sketch(const SketchTask<S>* d, const SketchTile* tiles, int ntiles, const S* om, int ldo, hipStream_t s);
// BF16 sketch (FP32 C and Y): Omega as bf16 in HBM (ldo a multiple of 8, at least the largest
// w rounded up to 32), one 64 x tile_cols tile per workgroup (tile_cols 64, 128 or 256; the
// tiles' c0 are multiples of tile_cols).  launch_gauss_bf16 writes the FP32 Omega of
// launch_gauss rounded to bf16.
void launch_gauss_bf16(void* om, int rows, int cols, int ld, uint64_t seed, hipStream_t s);
void launch_sketch_bf16(const SketchTask<float>* d, const SketchTile* tiles, int ntiles, const void* om, int ldo,
                        int tile_cols, hipStream_t s);
template <typename S> void launch_rank_features(const FeatTask<S>* d, int n, double tol, hipStream_t s);
template <typename S> void launch_tail_select(const TailTask<S>* d, int n, double tol, hipStream_t s);
template <typename S> void launch_tail_lead(const TailLeadTask<S>* d, int n, double tol, hipStream_t s);
void launch_tail_gemm(const TailGemmTask* d, const SketchTile* tiles, int ntiles, hipStream_t s);
template <typename S> void launch_eye(const EyeTask<S>* d, int n, hipStream_t s);

// MLP (kRankFeatures -> 64 -> 64 -> 1, tanh, sigmoid output): predicted rank fractions.
void launch_mlp_predict(const float* X, int n, const float* W, float* out, hipStream_t s);
// One full-batch gradient of the mean logistic loss; partial sums per workgroup are reduced
// in a fixed order, so the result is deterministic.
int mlp_grad_blocks(int n);
void launch_mlp_grad(const float* X, const float* y, int n, const float* W, float* partial, int nblocks,
                     hipStream_t s);
// grad = sum of partials (and loss); with adam != 0 also one Adam step on W (m, v moments).
void launch_mlp_reduce(const float* partial, int nblocks, float* grad, float* loss, float* W, float* m, float* v,
                       int step, float lr, int adam, hipStream_t s);

// Host-memory wrappers (allocate, copy, run, copy back, synchronize) for the C ABI.
int mlp_predict_host(const float* X, int n, const float* W, float* out);
int mlp_grad_host(const float* X, const float* y, int n, const float* W, float* grad, float* loss);
int mlp_train_host(const float* X, const float* y, int n, float* W, int epochs, float lr, float* loss_hist);

}  // namespace spaqr
