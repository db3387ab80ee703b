// The per-matrix symbolic phase of a square system on the device (symbolic.hpp).
#include "symbolic.hpp"

#include <cstring>  // rocprim/iterator/texture_cache_iterator.hpp calls memset

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

namespace spaqr {
namespace {

#define SCHK(x)                                                                                   \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) throw std::runtime_error(std::string("symbolic phase: ") + hipGetErrorString(e_)); \
    } while (0)

constexpr int TB = 256;

inline unsigned nblk(long long n) { return (unsigned)std::max<long long>(1, (n + TB - 1) / TB); }

template <class T>
T* dalloc(size_t n) {
    T* p = nullptr;
    SCHK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    return p;
}
template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

// rocPRIM temporary storage, grown as needed
struct Tmp {
    void* p = nullptr;
    size_t sz = 0;
    void* get(size_t need) {
        if (need > sz) {
            if (p) SCHK(hipFree(p));
            p = nullptr;
            SCHK(hipMalloc(&p, std::max<size_t>(need, 256)));
            sz = need;
        }
        return p;
    }
    ~Tmp() {
        if (p) (void)hipFree(p);
    }
};

// Keys are sorted as unsigned integers over their low `bits` bits (every key here is
// non-negative, so a signed key is sorted through its unsigned view).
template <class K> struct ukey { using type = K; };
template <> struct ukey<long long> { using type = unsigned long long; };
template <> struct ukey<int> { using type = unsigned; };
template <class K, class V>
void sort_pairs(Tmp& t, const K* kin, K* kout, const V* vin, V* vout, size_t n, unsigned bits, hipStream_t s) {
    using U = typename ukey<K>::type;
    const U* ki = reinterpret_cast<const U*>(kin);
    U* ko = reinterpret_cast<U*>(kout);
    size_t sz = 0;
    SCHK(rocprim::radix_sort_pairs(nullptr, sz, ki, ko, vin, vout, n, 0u, bits, s));
    SCHK(rocprim::radix_sort_pairs(t.get(sz), sz, ki, ko, vin, vout, n, 0u, bits, s));
}
template <class K>
void sort_keys(Tmp& t, const K* kin, K* kout, size_t n, unsigned bits, hipStream_t s) {
    using U = typename ukey<K>::type;
    const U* ki = reinterpret_cast<const U*>(kin);
    U* ko = reinterpret_cast<U*>(kout);
    size_t sz = 0;
    SCHK(rocprim::radix_sort_keys(nullptr, sz, ki, ko, n, 0u, bits, s));
    SCHK(rocprim::radix_sort_keys(t.get(sz), sz, ki, ko, n, 0u, bits, s));
}
template <class T>
void exscan(Tmp& t, const T* in, T* out, size_t n, hipStream_t s) {
    size_t sz = 0;
    SCHK(rocprim::exclusive_scan(nullptr, sz, in, out, T(0), n, rocprim::plus<T>(), s));
    SCHK(rocprim::exclusive_scan(t.get(sz), sz, in, out, T(0), n, rocprim::plus<T>(), s));
}
// sorted keys -> their distinct values (in place order), returns the count
size_t unique_keys(Tmp& t, const long long* in, long long* out, size_t n, hipStream_t s) {
    size_t* d_cnt = dalloc<size_t>(1);
    size_t sz = 0;
    SCHK(rocprim::unique(nullptr, sz, in, out, d_cnt, n, rocprim::equal_to<long long>(), s));
    SCHK(rocprim::unique(t.get(sz), sz, in, out, d_cnt, n, rocprim::equal_to<long long>(), s));
    size_t h = 0;
    SCHK(hipMemcpyAsync(&h, d_cnt, sizeof(size_t), hipMemcpyDeviceToHost, s));
    SCHK(hipStreamSynchronize(s));
    dfree(d_cnt);
    return h;
}

unsigned bits_for(unsigned long long v) {
    unsigned b = 1;
    while (b < 64 && (v >> b) != 0) ++b;
    return b;
}

// order-preserving map of a double onto unsigned 64-bit integers
__device__ __forceinline__ unsigned long long enc_double(double x) {
    const unsigned long long b = __double_as_longlong(x);
    return (b >> 63) ? ~b : (b | (1ull << 63));
}
__device__ __forceinline__ double dec_double(unsigned long long e) {
    const unsigned long long b = (e >> 63) ? (e & ~(1ull << 63)) : ~e;
    return __longlong_as_double((long long)b);
}

// ---- A's CSR and the A^T A pattern
__global__ void k_csc_keys(int n, const int* __restrict__ cp, const int* __restrict__ ri, long long* __restrict__ key,
                           int* __restrict__ rowcnt) {
    const int j = blockIdx.x * TB + threadIdx.x;
    if (j >= n) return;
    for (int k = cp[j]; k < cp[j + 1]; ++k) {
        key[k] = (long long)ri[k] * n + j;
        atomicAdd(&rowcnt[ri[k]], 1);
    }
}
__global__ void k_key_col(long long nk, int n, const long long* __restrict__ key, int* __restrict__ col) {
    const long long q = (long long)blockIdx.x * TB + threadIdx.x;
    if (q < nk) col[q] = (int)(key[q] % n);
}
__global__ void k_ata_count(int n, const int* __restrict__ cp, const int* __restrict__ ri, const int* __restrict__ rp,
                            long long* __restrict__ cnt) {
    const int j = blockIdx.x * TB + threadIdx.x;
    if (j >= n) return;
    long long c = 0;
    for (int k = cp[j]; k < cp[j + 1]; ++k) c += rp[ri[k] + 1] - rp[ri[k]];
    cnt[j] = c;
}
__global__ void k_ata_write(int n, const int* __restrict__ cp, const int* __restrict__ ri, const int* __restrict__ rp,
                            const int* __restrict__ ci, const long long* __restrict__ off, long long* __restrict__ key) {
    const int j = blockIdx.x * TB + threadIdx.x;
    if (j >= n) return;
    long long o = off[j];
    for (int k = cp[j]; k < cp[j + 1]; ++k) {
        const int r = ri[k];
        for (int q = rp[r]; q < rp[r + 1]; ++q) key[o++] = (long long)j * n + ci[q];
    }
}
__global__ void k_key_rowcount(long long nk, int n, const long long* __restrict__ key, int* __restrict__ cnt) {
    const long long q = (long long)blockIdx.x * TB + threadIdx.x;
    if (q < nk) atomicAdd(&cnt[key[q] / n], 1);
}

// ---- geometric bisection (bissect_geo, src/partition.cpp:241-279)
// Per part: the vertex count and each dimension's range.  One pass per distinct part among a
// wave's vertices (consecutive vertices meet few parts): that part's lanes reduce count, min
// and max over the wave and one lane issues the atomics, so a part's counters see one update
// per wave, not one per vertex.
__global__ void k_minmax(int n, int d, const double* __restrict__ X, const int* __restrict__ parts,
                         unsigned long long* __restrict__ mm, int* __restrict__ cnt) {
    const int i = blockIdx.x * TB + threadIdx.x;
    const bool live = i < n;
    const int k = live ? parts[i] : -1;
    const int lane = threadIdx.x & 63;
    unsigned long long left = __ballot(live && k >= 0);
    while (left) {
        const int leader = __ffsll((long long)left) - 1;
        const int kk = __shfl(k, leader, 64);
        const bool mine = live && k == kk;
        const unsigned long long grp = __ballot(mine);
        left &= ~grp;
        if (lane == leader) atomicAdd(&cnt[kk], __popcll(grp));
        for (int dim = 0; dim < d; ++dim) {
            const unsigned long long e = mine ? enc_double(X[dim + (size_t)i * d]) : 0ull;
            unsigned long long lo = mine ? e : ~0ull, hi = mine ? e : 0ull;
            for (int o = 32; o > 0; o >>= 1) {
                lo = min(lo, __shfl_xor(lo, o, 64));
                hi = max(hi, __shfl_xor(hi, o, 64));
            }
            if (lane == leader) {
                atomicMin(&mm[2 * ((size_t)kk * d + dim)], lo);
                atomicMax(&mm[2 * ((size_t)kk * d + dim) + 1], hi);
            }
        }
    }
}
// the widest dimension of every part: the first whose range is strictly the largest
__global__ void k_bestdim(int np, int d, const unsigned long long* __restrict__ mm, const int* __restrict__ cnt,
                          int* __restrict__ best) {
    const int k = blockIdx.x * TB + threadIdx.x;
    if (k >= np) return;
    int b = 0;
    double range = -1.0;
    if (cnt[k] > 0)
        for (int dim = 0; dim < d; ++dim) {
            const double lo = dec_double(mm[2 * ((size_t)k * d + dim)]), hi = dec_double(mm[2 * ((size_t)k * d + dim) + 1]);
            if (hi - lo > range) { range = hi - lo; b = dim; }
        }
    best[k] = b;
}
__global__ void k_coord_keys(int n, int d, const double* __restrict__ X, const int* __restrict__ parts,
                             const int* __restrict__ best, unsigned long long* __restrict__ key, int* __restrict__ idx) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    key[i] = enc_double(X[best[parts[i]] + (size_t)i * d]);
    idx[i] = i;
}
__global__ void k_part_of(int n, const int* __restrict__ idx, const int* __restrict__ parts, unsigned* __restrict__ key) {
    const int q = blockIdx.x * TB + threadIdx.x;
    if (q < n) key[q] = (unsigned)parts[idx[q]];
}
// the (N/2)-th smallest coordinate of every nonempty part (the value decides, <= goes left)
__global__ void k_mid(int np, int d, const double* __restrict__ X, const int* __restrict__ idx, const int* __restrict__ start,
                      const int* __restrict__ cnt, const int* __restrict__ best, double* __restrict__ mid) {
    const int k = blockIdx.x * TB + threadIdx.x;
    if (k >= np || cnt[k] == 0) return;
    mid[k] = X[best[k] + (size_t)idx[start[k] + cnt[k] / 2] * d];
}
__global__ void k_assign(int n, int d, const double* __restrict__ X, const int* __restrict__ best,
                         const double* __restrict__ mid, int* __restrict__ parts) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    const int k = parts[i];
    parts[i] = (X[best[k] + (size_t)i * d] <= mid[k]) ? 2 * k : 2 * k + 1;
}

// ---- separators (src/partition.cpp:306-333): a left-half vertex with an A^T A neighbour in the
// right half of the same part, neither yet a separator.  Only left-half vertices change in one
// depth, so the right-half test reads values that no thread of the launch writes.
__global__ void k_separators(int n, int l, const int* __restrict__ gp, const int* __restrict__ gi,
                             const int* __restrict__ parts, int* __restrict__ slvl, int* __restrict__ ssep,
                             int* __restrict__ section) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n || slvl[i] != -1) return;
    const int sh = l - 1, pi = parts[i] >> sh;
    if (pi & 1) return;
    bool sep = false;
    for (int k = gp[i]; k < gp[i + 1] && !sep; ++k) {
        const int j = gi[k];
        if ((parts[j] >> sh) == pi + 1 && slvl[j] == -1) sep = true;
    }
    if (sep) {
        slvl[i] = l;
        ssep[i] = pi / 2;
        section[i] = parts[i];
    }
}
__global__ void k_leaves(int n, const int* __restrict__ parts, int* slvl, int* ssep, int* llvl, int* lsep, int* rlvl,
                         int* rsep, int* section) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    if (slvl[i] == -1) {
        slvl[i] = llvl[i] = rlvl[i] = 0;
        ssep[i] = lsep[i] = rsep[i] = parts[i];
        section[i] = parts[i];
    } else {
        llvl[i] = rlvl[i] = -1;
        lsep[i] = rsep[i] = 0;
    }
}

struct Sep {
    int lvl, sep;
};
__device__ __forceinline__ Sep lowest_common_d(Sep a, Sep b) { return a.lvl < b.lvl ? a : b; }
__device__ __forceinline__ Sep highest_common_d(Sep a, Sep b) {
    while (a.lvl < b.lvl) { a.lvl++; a.sep /= 2; }
    while (b.lvl < a.lvl) { b.lvl++; b.sep /= 2; }
    while (a.sep != b.sep) { a.lvl++; b.lvl++; a.sep /= 2; b.sep /= 2; }
    return a;
}

// getInterfacesUsingA (src/partition.cpp:514-587), rmap == cmap: first pass -- a separator
// vertex's left and right neighbours among the lower levels along its row of A, in column order
__global__ void k_iface1(int n, const int* __restrict__ rp, const int* __restrict__ ci, const int* __restrict__ slvl,
                         const int* __restrict__ ssep, int* __restrict__ llvl, int* __restrict__ lsep,
                         int* __restrict__ rlvl, int* __restrict__ rsep) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n || slvl[i] <= 0) return;
    const int sl = slvl[i];
    Sep left{-1, 0}, right{-1, 0};
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
        const int j = ci[k];
        const Sep nb{slvl[j], ssep[j]};
        if (nb.lvl < sl) {
            const int pj = nb.sep >> (sl - nb.lvl - 1);
            if ((pj & 1) == 0) left = (left.lvl == -1) ? nb : lowest_common_d(nb, left);
            else right = (right.lvl == -1) ? nb : lowest_common_d(nb, right);
        }
    }
    llvl[i] = left.lvl; lsep[i] = left.sep;
    rlvl[i] = right.lvl; rsep[i] = right.sep;
}
// second pass: a missing side from the row's vertices of the same separator (highest common
// ancestor), reading the first pass's values (in) and writing out; third: still missing -> self
__global__ void k_iface2(int n, const int* __restrict__ rp, const int* __restrict__ ci, const int* __restrict__ slvl,
                         const int* __restrict__ ssep, const int* __restrict__ li, const int* __restrict__ lsi,
                         const int* __restrict__ ri, const int* __restrict__ rsi, int* __restrict__ lo,
                         int* __restrict__ lso, int* __restrict__ ro, int* __restrict__ rso) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    Sep left{li[i], lsi[i]}, right{ri[i], rsi[i]};
    const int sl = slvl[i], ss = ssep[i];
    if (right.lvl == -1 && right.sep == 0) {
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            const int j = ci[k];
            if (slvl[j] == sl && ssep[j] == ss && !(ri[j] == -1 && rsi[j] == 0)) {
                const Sep cr{ri[j], rsi[j]};
                right = (right.lvl == -1 && right.sep == 0) ? cr : highest_common_d(right, cr);
            }
        }
    } else if (left.lvl == -1 && left.sep == 0) {
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            const int j = ci[k];
            if (slvl[j] == sl && ssep[j] == ss && !(li[j] == -1 && lsi[j] == 0)) {
                const Sep cl{li[j], lsi[j]};
                left = (left.lvl == -1 && left.sep == 0) ? cl : highest_common_d(left, cl);
            }
        }
    }
    if (left.lvl == -1 && left.sep == 0) left = Sep{sl, ss};
    if (right.lvl == -1 && right.sep == 0) right = Sep{sl, ss};
    lo[i] = left.lvl; lso[i] = left.sep;
    ro[i] = right.lvl; rso[i] = right.sep;
}

// ---- distinct IDs and the permutation
__global__ void k_id_keys(int n, const int* slvl, const int* ssep, const int* llvl, const int* lsep, const int* rlvl,
                          const int* rsep, const int* section, unsigned long long* hi, unsigned long long* lo,
                          int* idx) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    hi[i] = ((unsigned long long)(slvl[i] + 1) << 37) | ((unsigned long long)ssep[i] << 21) |
            ((unsigned long long)(llvl[i] + 1) << 16) | (unsigned long long)lsep[i];
    lo[i] = ((unsigned long long)(rlvl[i] + 1) << 32) | ((unsigned long long)rsep[i] << 16) | (unsigned long long)section[i];
    idx[i] = i;
}
__global__ void k_gather_u64(int n, const int* __restrict__ idx, const unsigned long long* __restrict__ src,
                             unsigned long long* __restrict__ dst) {
    const int q = blockIdx.x * TB + threadIdx.x;
    if (q < n) dst[q] = src[idx[q]];
}
__global__ void k_id_flags(int n, const unsigned long long* hi, const unsigned long long* lo, int* flag) {
    const int q = blockIdx.x * TB + threadIdx.x;
    if (q >= n) return;
    flag[q] = (q == 0 || hi[q] != hi[q - 1] || lo[q] != lo[q - 1]) ? 1 : 0;
}
__global__ void k_id_scatter(int n, const int* __restrict__ idx, const int* __restrict__ flag, const int* __restrict__ pos,
                             const unsigned long long* __restrict__ hi, const unsigned long long* __restrict__ lo,
                             int* __restrict__ did, unsigned long long* __restrict__ uid) {
    const int q = blockIdx.x * TB + threadIdx.x;
    if (q >= n) return;
    const int dq = pos[q] + flag[q] - 1;  // exclusive scan of the flags
    did[idx[q]] = dq;
    if (flag[q]) { uid[2 * dq] = hi[q]; uid[2 * dq + 1] = lo[q]; }
}
__global__ void k_rank_keys(int n, const int* __restrict__ did, const int* __restrict__ rank, unsigned* __restrict__ key,
                            int* __restrict__ idx) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    key[i] = (unsigned)rank[did[i]];
    idx[i] = i;
}

// ---- assembly
__global__ void k_inverse(int n, const int* __restrict__ perm, int* __restrict__ inv) {
    const int k = blockIdx.x * TB + threadIdx.x;
    if (k < n) inv[perm[k]] = k;
}
__global__ void k_edge_keys(int n, int n0, const int* __restrict__ cp, const int* __restrict__ ri,
                            const int* __restrict__ pinv, const int* __restrict__ owner, long long* __restrict__ key) {
    const int j = blockIdx.x * TB + threadIdx.x;
    if (j >= n) return;
    const long long s = owner[pinv[j]];
    for (int k = cp[j]; k < cp[j + 1]; ++k) key[k] = s * n0 + owner[pinv[ri[k]]];
}
__global__ void k_self_keys(int n0, long long* __restrict__ key) {
    const int s = blockIdx.x * TB + threadIdx.x;
    if (s < n0) key[s] = (long long)s * n0 + s;
}
template <typename S>
__global__ void k_scatter_blocks(int n, int n0, const int* __restrict__ cp, const int* __restrict__ ri,
                                 const double* __restrict__ v, const int* __restrict__ pinv, const int* __restrict__ owner,
                                 const long long* __restrict__ ekeys, int ne, const long long* __restrict__ off,
                                 const int* __restrict__ cstart, const int* __restrict__ cols, S* __restrict__ dst) {
    const int j = blockIdx.x * TB + threadIdx.x;
    if (j >= n) return;
    const int pc = pinv[j], s = owner[pc];
    for (int k = cp[j]; k < cp[j + 1]; ++k) {
        const int pr = pinv[ri[k]], o = owner[pr];
        const long long key = (long long)s * n0 + o;
        int lo = 0, hi = ne;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (ekeys[mid] < key) lo = mid + 1; else hi = mid;
        }
        dst[off[lo] + (pr - cstart[o]) + (size_t)(pc - cstart[s]) * cols[o]] = (S)v[k];
    }
}
__global__ void k_app_keys(int n, const int* __restrict__ cp, const int* __restrict__ ri, const int* __restrict__ pinv,
                           long long* __restrict__ key, int* __restrict__ rowcnt) {
    const int j = blockIdx.x * TB + threadIdx.x;
    if (j >= n) return;
    const long long pc = pinv[j];
    for (int k = cp[j]; k < cp[j + 1]; ++k) {
        const int pr = pinv[ri[k]];
        key[k] = (long long)pr * n + pc;
        atomicAdd(&rowcnt[pr], 1);
    }
}
__global__ void k_pair_keys(int n, int n0, const int* __restrict__ gp, const int* __restrict__ gi,
                            const int* __restrict__ pinv, const int* __restrict__ owner, long long* __restrict__ key) {
    const int j = blockIdx.x * TB + threadIdx.x;
    if (j >= n) return;
    const long long s = owner[pinv[j]];
    for (int k = gp[j]; k < gp[j + 1]; ++k) key[k] = s * n0 + owner[pinv[gi[k]]];
}

}  // namespace

void SymbolicDevice::release() {
    dfree(d_cp); dfree(d_ri); dfree(d_v); dfree(d_rp); dfree(d_ci); dfree(d_gp); dfree(d_gi);
    dfree(d_pinv); dfree(d_owner); dfree(d_ekeys);
    n = nnz = n0 = ne = 0;
    gnnz = 0;
}

void SymbolicDevice::order(int n_, const int* colptr, const int* rowind, const double* val, const double* X, int xdim,
                           int nlevels, const std::function<std::vector<int>(const std::vector<DevId>&)>& rank_ids,
                           std::vector<int>& cperm, std::vector<int>& did, std::vector<DevId>& ids,
                           std::vector<int>& parts, hipStream_t s) {
    static const bool dbg = getenv("SPAQR_DEBUG_HOST") != nullptr;
    auto tp = std::chrono::steady_clock::now();
    auto lap = [&](const char* what) {
        if (!dbg) return;
        SCHK(hipStreamSynchronize(s));
        const auto nw = std::chrono::steady_clock::now();
        fprintf(stderr, "[partition/device] %s %.1f ms\n", what, std::chrono::duration<double, std::milli>(nw - tp).count());
        tp = nw;
    };
    release();
    n = n_;
    nnz = colptr[n];
    if (nlevels > 17) throw std::runtime_error("the device ordering handles at most 17 levels");
    Tmp tmp;
    d_cp = dalloc<int>(n + 1);
    d_ri = dalloc<int>(nnz);
    d_v = dalloc<double>(nnz);
    SCHK(hipMemcpyAsync(d_cp, colptr, sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    SCHK(hipMemcpyAsync(d_ri, rowind, sizeof(int) * nnz, hipMemcpyHostToDevice, s));
    SCHK(hipMemcpyAsync(d_v, val, sizeof(double) * nnz, hipMemcpyHostToDevice, s));
    // A's CSR (row i's columns in increasing order)
    {
        long long* k0 = dalloc<long long>(nnz);
        long long* k1 = dalloc<long long>(nnz);
        int* cnt = dalloc<int>(n + 1);
        SCHK(hipMemsetAsync(cnt, 0, sizeof(int) * (n + 1), s));
        hipLaunchKernelGGL(k_csc_keys, dim3(nblk(n)), dim3(TB), 0, s, n, d_cp, d_ri, k0, cnt);
        sort_keys(tmp, k0, k1, (size_t)nnz, bits_for((unsigned long long)n * n), s);
        d_rp = dalloc<int>(n + 1);
        d_ci = dalloc<int>(nnz);
        exscan(tmp, cnt, d_rp, (size_t)n + 1, s);
        hipLaunchKernelGGL(k_key_col, dim3(nblk(nnz)), dim3(TB), 0, s, (long long)nnz, n, k1, d_ci);
        dfree(k0); dfree(k1); dfree(cnt);
    }
    lap("upload + CSR");
    // A^T A pattern: candidates per column, sorted, distinct
    {
        long long* cnt = dalloc<long long>(n + 1);
        long long* off = dalloc<long long>(n + 1);
        SCHK(hipMemsetAsync(cnt, 0, sizeof(long long) * (n + 1), s));
        hipLaunchKernelGGL(k_ata_count, dim3(nblk(n)), dim3(TB), 0, s, n, d_cp, d_ri, d_rp, cnt);
        exscan(tmp, cnt, off, (size_t)n + 1, s);
        long long total = 0;
        SCHK(hipMemcpyAsync(&total, off + n, sizeof(long long), hipMemcpyDeviceToHost, s));
        SCHK(hipStreamSynchronize(s));
        long long* k0 = dalloc<long long>(total);
        long long* k1 = dalloc<long long>(total);
        hipLaunchKernelGGL(k_ata_write, dim3(nblk(n)), dim3(TB), 0, s, n, d_cp, d_ri, d_rp, d_ci, off, k0);
        sort_keys(tmp, k0, k1, (size_t)total, bits_for((unsigned long long)n * n), s);
        gnnz = (long long)unique_keys(tmp, k1, k0, (size_t)total, s);
        int* gcnt = dalloc<int>(n + 1);
        SCHK(hipMemsetAsync(gcnt, 0, sizeof(int) * (n + 1), s));
        hipLaunchKernelGGL(k_key_rowcount, dim3(nblk(gnnz)), dim3(TB), 0, s, gnnz, n, k0, gcnt);
        d_gp = dalloc<int>(n + 1);
        d_gi = dalloc<int>(gnnz);
        exscan(tmp, gcnt, d_gp, (size_t)n + 1, s);
        hipLaunchKernelGGL(k_key_col, dim3(nblk(gnnz)), dim3(TB), 0, s, gnnz, n, k0, d_gi);
        dfree(k0); dfree(k1); dfree(cnt); dfree(off); dfree(gcnt);
    }
    lap("A^T A pattern");
    // recursive geometric bisection
    int* d_parts = dalloc<int>(n);
    SCHK(hipMemsetAsync(d_parts, 0, sizeof(int) * n, s));
    {
        double* d_X = dalloc<double>((size_t)xdim * n);
        SCHK(hipMemcpyAsync(d_X, X, sizeof(double) * (size_t)xdim * n, hipMemcpyHostToDevice, s));
        const int npmax = 1 << std::max(nlevels - 2, 0);
        unsigned long long* mm = dalloc<unsigned long long>(2 * (size_t)npmax * xdim);
        int *cnt = dalloc<int>(npmax + 1), *start = dalloc<int>(npmax + 1), *best = dalloc<int>(npmax);
        double* mid = dalloc<double>(npmax);
        unsigned long long *k0 = dalloc<unsigned long long>(n), *k1 = dalloc<unsigned long long>(n);
        unsigned *p0 = dalloc<unsigned>(n), *p1 = dalloc<unsigned>(n);
        int *i0 = dalloc<int>(n), *i1 = dalloc<int>(n);
        std::vector<unsigned long long> init(2 * (size_t)npmax * xdim);
        for (size_t q = 0; q < init.size(); ++q) init[q] = (q & 1) ? 0ull : ~0ull;
        for (int depth = 0; depth < nlevels - 1; ++depth) {
            const int np = 1 << depth;
            SCHK(hipMemcpyAsync(mm, init.data(), sizeof(unsigned long long) * 2 * (size_t)np * xdim, hipMemcpyHostToDevice, s));
            SCHK(hipMemsetAsync(cnt, 0, sizeof(int) * (np + 1), s));
            hipLaunchKernelGGL(k_minmax, dim3(nblk(n)), dim3(TB), 0, s, n, xdim, d_X, d_parts, mm, cnt);
            hipLaunchKernelGGL(k_bestdim, dim3(nblk(np)), dim3(TB), 0, s, np, xdim, mm, cnt, best);
            hipLaunchKernelGGL(k_coord_keys, dim3(nblk(n)), dim3(TB), 0, s, n, xdim, d_X, d_parts, best, k0, i0);
            sort_pairs(tmp, k0, k1, i0, i1, (size_t)n, 64u, s);
            hipLaunchKernelGGL(k_part_of, dim3(nblk(n)), dim3(TB), 0, s, n, i1, d_parts, p0);
            sort_pairs(tmp, p0, p1, i1, i0, (size_t)n, (unsigned)std::max(depth, 1), s);
            exscan(tmp, cnt, start, (size_t)np + 1, s);
            hipLaunchKernelGGL(k_mid, dim3(nblk(np)), dim3(TB), 0, s, np, xdim, d_X, i0, start, cnt, best, mid);
            hipLaunchKernelGGL(k_assign, dim3(nblk(n)), dim3(TB), 0, s, n, xdim, d_X, best, mid, d_parts);
        }
        SCHK(hipStreamSynchronize(s));  // init (host) was the source of asynchronous copies
        dfree(d_X); dfree(mm); dfree(cnt); dfree(start); dfree(best); dfree(mid);
        dfree(k0); dfree(k1); dfree(p0); dfree(p1); dfree(i0); dfree(i1);
    }
    lap("bisection");
    // separators, leaves, interfaces
    int *slvl = dalloc<int>(n), *ssep = dalloc<int>(n), *llvl = dalloc<int>(n), *lsep = dalloc<int>(n);
    int *rlvl = dalloc<int>(n), *rsep = dalloc<int>(n), *sect = dalloc<int>(n);
    int *l2 = dalloc<int>(n), *ls2 = dalloc<int>(n), *r2 = dalloc<int>(n), *rs2 = dalloc<int>(n);
    {
        std::vector<int> m1(n, -1);
        SCHK(hipMemcpyAsync(slvl, m1.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
        SCHK(hipMemsetAsync(ssep, 0, sizeof(int) * n, s));
        SCHK(hipMemsetAsync(sect, 0, sizeof(int) * n, s));
        for (int depth = 0; depth < nlevels - 1; ++depth)
            hipLaunchKernelGGL(k_separators, dim3(nblk(n)), dim3(TB), 0, s, n, nlevels - depth - 1, d_gp, d_gi, d_parts,
                               slvl, ssep, sect);
        hipLaunchKernelGGL(k_leaves, dim3(nblk(n)), dim3(TB), 0, s, n, d_parts, slvl, ssep, llvl, lsep, rlvl, rsep, sect);
        hipLaunchKernelGGL(k_iface1, dim3(nblk(n)), dim3(TB), 0, s, n, d_rp, d_ci, slvl, ssep, llvl, lsep, rlvl, rsep);
        hipLaunchKernelGGL(k_iface2, dim3(nblk(n)), dim3(TB), 0, s, n, d_rp, d_ci, slvl, ssep, llvl, lsep, rlvl, rsep,
                           l2, ls2, r2, rs2);
        SCHK(hipStreamSynchronize(s));  // m1 (host) was the source of an asynchronous copy
    }
    lap("separators + interfaces");
    // distinct IDs, their ranks, the permutation
    {
        unsigned long long *hi = dalloc<unsigned long long>(n), *lo = dalloc<unsigned long long>(n);
        unsigned long long *t0 = dalloc<unsigned long long>(n), *t1 = dalloc<unsigned long long>(n);
        int *i0 = dalloc<int>(n), *i1 = dalloc<int>(n), *flag = dalloc<int>(n), *pos = dalloc<int>(n);
        int* d_did = dalloc<int>(n);
        hipLaunchKernelGGL(k_id_keys, dim3(nblk(n)), dim3(TB), 0, s, n, slvl, ssep, l2, ls2, r2, rs2, sect, hi, lo, i0);
        sort_pairs(tmp, lo, t0, i0, i1, (size_t)n, 37u, s);
        hipLaunchKernelGGL(k_gather_u64, dim3(nblk(n)), dim3(TB), 0, s, n, i1, hi, t0);
        sort_pairs(tmp, t0, t1, i1, i0, (size_t)n, 42u, s);  // i0: vertices by (hi, lo)
        hipLaunchKernelGGL(k_gather_u64, dim3(nblk(n)), dim3(TB), 0, s, n, i0, lo, t0);
        hipLaunchKernelGGL(k_id_flags, dim3(nblk(n)), dim3(TB), 0, s, n, t1, t0, flag);
        exscan(tmp, flag, pos, (size_t)n, s);
        int nd = 0, lastf = 0;
        SCHK(hipMemcpyAsync(&nd, pos + n - 1, sizeof(int), hipMemcpyDeviceToHost, s));
        SCHK(hipMemcpyAsync(&lastf, flag + n - 1, sizeof(int), hipMemcpyDeviceToHost, s));
        SCHK(hipStreamSynchronize(s));
        nd += lastf;
        unsigned long long* uid = dalloc<unsigned long long>(2 * (size_t)nd);
        hipLaunchKernelGGL(k_id_scatter, dim3(nblk(n)), dim3(TB), 0, s, n, i0, flag, pos, t1, t0, d_did, uid);
        std::vector<unsigned long long> h(2 * (size_t)nd);
        SCHK(hipMemcpyAsync(h.data(), uid, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost, s));
        SCHK(hipStreamSynchronize(s));
        ids.resize(nd);
        for (int q = 0; q < nd; ++q) {
            const unsigned long long a = h[2 * q], b = h[2 * q + 1];
            DevId& d = ids[q];
            d.slvl = (int)(a >> 37) - 1;
            d.ssep = (int)((a >> 21) & 0xffff);
            d.llvl = (int)((a >> 16) & 0x1f) - 1;
            d.lsep = (int)(a & 0xffff);
            d.rlvl = (int)(b >> 32) - 1;
            d.rsep = (int)((b >> 16) & 0xffff);
            d.section = (int)(b & 0xffff);
        }
        lap("distinct IDs");
        const std::vector<int> rank = rank_ids(ids);
        lap("ranks (host)");
        int* d_rank = dalloc<int>(nd);
        SCHK(hipMemcpyAsync(d_rank, rank.data(), sizeof(int) * nd, hipMemcpyHostToDevice, s));
        unsigned *k0 = dalloc<unsigned>(n), *k1 = dalloc<unsigned>(n);
        hipLaunchKernelGGL(k_rank_keys, dim3(nblk(n)), dim3(TB), 0, s, n, d_did, d_rank, k0, i0);
        sort_pairs(tmp, k0, k1, i0, i1, (size_t)n, bits_for((unsigned long long)nd), s);
        cperm.resize(n);
        did.resize(n);
        parts.resize(n);
        SCHK(hipMemcpyAsync(cperm.data(), i1, sizeof(int) * n, hipMemcpyDeviceToHost, s));
        SCHK(hipMemcpyAsync(did.data(), d_did, sizeof(int) * n, hipMemcpyDeviceToHost, s));
        SCHK(hipMemcpyAsync(parts.data(), d_parts, sizeof(int) * n, hipMemcpyDeviceToHost, s));
        SCHK(hipStreamSynchronize(s));
        dfree(hi); dfree(lo); dfree(t0); dfree(t1); dfree(i0); dfree(i1); dfree(flag); dfree(pos); dfree(d_did);
        dfree(uid); dfree(d_rank); dfree(k0); dfree(k1);
    }
    lap("permutation");
    dfree(d_parts);
    dfree(slvl); dfree(ssep); dfree(llvl); dfree(lsep); dfree(rlvl); dfree(rsep); dfree(sect);
    dfree(l2); dfree(ls2); dfree(r2); dfree(rs2);
}

void SymbolicDevice::edges(const std::vector<int>& cperm, const std::vector<int>& owner, int n0_,
                           std::vector<long long>& keys, hipStream_t s) {
    if (!d_cp) throw std::runtime_error("the device symbolic phase has no matrix (partition first)");
    n0 = n0_;
    Tmp tmp;
    int* d_perm = dalloc<int>(n);
    dfree(d_pinv);
    dfree(d_owner);
    d_pinv = dalloc<int>(n);
    d_owner = dalloc<int>(n);
    SCHK(hipMemcpyAsync(d_perm, cperm.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    SCHK(hipMemcpyAsync(d_owner, owner.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_inverse, dim3(nblk(n)), dim3(TB), 0, s, n, d_perm, d_pinv);
    long long* k0 = dalloc<long long>((size_t)nnz + n0);
    long long* k1 = dalloc<long long>((size_t)nnz + n0);
    hipLaunchKernelGGL(k_edge_keys, dim3(nblk(n)), dim3(TB), 0, s, n, n0, d_cp, d_ri, d_pinv, d_owner, k0);
    hipLaunchKernelGGL(k_self_keys, dim3(nblk(n0)), dim3(TB), 0, s, n0, k0 + nnz);
    sort_keys(tmp, k0, k1, (size_t)nnz + n0, bits_for((unsigned long long)n0 * n0), s);
    ne = (int)unique_keys(tmp, k1, k0, (size_t)nnz + n0, s);
    dfree(d_ekeys);
    d_ekeys = dalloc<long long>(ne);
    SCHK(hipMemcpyAsync(d_ekeys, k0, sizeof(long long) * ne, hipMemcpyDeviceToDevice, s));
    keys.resize(ne);
    SCHK(hipMemcpyAsync(keys.data(), k0, sizeof(long long) * ne, hipMemcpyDeviceToHost, s));
    SCHK(hipStreamSynchronize(s));
    dfree(k0); dfree(k1); dfree(d_perm);
}

template <typename S>
void SymbolicDevice::scatter_blocks(const std::vector<long long>& off, const std::vector<int>& cstart,
                                    const std::vector<int>& cols, S* dst, size_t nel, hipStream_t s) {
    if ((int)off.size() != ne || (int)cstart.size() != n0) throw std::runtime_error("block table size mismatch");
    long long* d_off = dalloc<long long>(ne);
    int *d_cs = dalloc<int>(n0), *d_cols = dalloc<int>(n0);
    SCHK(hipMemcpyAsync(d_off, off.data(), sizeof(long long) * ne, hipMemcpyHostToDevice, s));
    SCHK(hipMemcpyAsync(d_cs, cstart.data(), sizeof(int) * n0, hipMemcpyHostToDevice, s));
    SCHK(hipMemcpyAsync(d_cols, cols.data(), sizeof(int) * n0, hipMemcpyHostToDevice, s));
    SCHK(hipMemsetAsync(dst, 0, sizeof(S) * std::max<size_t>(nel, 1), s));
    hipLaunchKernelGGL(k_scatter_blocks<S>, dim3(nblk(n)), dim3(TB), 0, s, n, n0, d_cp, d_ri, d_v, d_pinv, d_owner,
                       d_ekeys, ne, d_off, d_cs, d_cols, dst);
    SCHK(hipStreamSynchronize(s));
    dfree(d_off); dfree(d_cs); dfree(d_cols);
}
template void SymbolicDevice::scatter_blocks<float>(const std::vector<long long>&, const std::vector<int>&,
                                                    const std::vector<int>&, float*, size_t, hipStream_t);
template void SymbolicDevice::scatter_blocks<double>(const std::vector<long long>&, const std::vector<int>&,
                                                     const std::vector<int>&, double*, size_t, hipStream_t);

void SymbolicDevice::csr(int** rowptr, int** col, double** val, hipStream_t s) {
    Tmp tmp;
    long long *k0 = dalloc<long long>(nnz), *k1 = dalloc<long long>(nnz);
    double* v1 = dalloc<double>(nnz);
    int* cnt = dalloc<int>(n + 1);
    SCHK(hipMemsetAsync(cnt, 0, sizeof(int) * (n + 1), s));
    hipLaunchKernelGGL(k_app_keys, dim3(nblk(n)), dim3(TB), 0, s, n, d_cp, d_ri, d_pinv, k0, cnt);
    // the values in key order: the CSC value array follows the keys of k_app_keys entry by entry
    sort_pairs(tmp, k0, k1, d_v, v1, (size_t)nnz, bits_for((unsigned long long)n * n), s);
    *rowptr = dalloc<int>(n + 1);
    *col = dalloc<int>(nnz);
    *val = v1;
    exscan(tmp, cnt, *rowptr, (size_t)n + 1, s);
    hipLaunchKernelGGL(k_key_col, dim3(nblk(nnz)), dim3(TB), 0, s, (long long)nnz, n, k1, *col);
    SCHK(hipStreamSynchronize(s));
    dfree(k0); dfree(k1); dfree(cnt);
}

void SymbolicDevice::ata_pairs(std::vector<long long>& keys, hipStream_t s) {
    Tmp tmp;
    long long *k0 = dalloc<long long>(gnnz), *k1 = dalloc<long long>(gnnz);
    hipLaunchKernelGGL(k_pair_keys, dim3(nblk(n)), dim3(TB), 0, s, n, n0, d_gp, d_gi, d_pinv, d_owner, k0);
    sort_keys(tmp, k0, k1, (size_t)gnnz, bits_for((unsigned long long)n0 * n0), s);
    const size_t np = unique_keys(tmp, k1, k0, (size_t)gnnz, s);
    keys.resize(np);
    SCHK(hipMemcpyAsync(keys.data(), k0, sizeof(long long) * np, hipMemcpyDeviceToHost, s));
    SCHK(hipStreamSynchronize(s));
    dfree(k0); dfree(k1);
}

}  // namespace spaqr
