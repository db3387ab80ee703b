// The per-matrix symbolic phase of a square system on the device: the geometric nested
// dissection (GeometricPartitionAtA + getInterfacesUsingA, src/partition.cpp:241-350,514-587),
// initPermutation (src/tree.cpp:48-57) and ParTree::assemble's block extraction, SpMV storage
// and distance-two pairs (src/ptree.cpp:696-864).  The O(nnz) work (A^T A pattern, bisection
// sorts, separator and interface rules, permutation sort, block scatter) runs in kernels and
// rocPRIM sorts; the cluster-level steps whose results depend on an order (the ranks of the
// distinct cluster IDs, the edge bookkeeping, the distance-two rule) stay on the host, over
// data that is already reduced to one entry per cluster or cluster pair.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <vector>

namespace spaqr {

// A level-0 cluster ID in plain ints (lvl -1 .. 15; sep and section below 2^16).
struct DevId {
    int slvl, ssep, llvl, lsep, rlvl, rsep, section;
};

class SymbolicDevice {
public:
    SymbolicDevice() {}
    ~SymbolicDevice() { release(); }
    SymbolicDevice(const SymbolicDevice&) = delete;
    SymbolicDevice& operator=(const SymbolicDevice&) = delete;

    // Ordering.  A is n x n CSC (host), X d x n coordinates.  rank_ids receives the distinct
    // level-0 IDs (in an arbitrary fixed order) and returns each one's position in the order
    // initPermutation's successive stable sorts give.  Outputs: cperm (natural index of permuted
    // position k), did (index into ids of every natural vertex), the leaf parts.
    void order(int n, const int* colptr, const int* rowind, const double* val, const double* X, int xdim, int nlevels,
               const std::function<std::vector<int>(const std::vector<DevId>&)>& rank_ids, std::vector<int>& cperm,
               std::vector<int>& did, std::vector<DevId>& ids, std::vector<int>& parts, hipStream_t s);

    // Assembly over the level-0 clusters (n0 of them; cluster of every permuted index in
    // owner, host; cstart, cols per cluster).  edges: every (self, nbr) pair with a nonzero of
    // App = A(cperm, cperm) in column block self and row block nbr, plus (self, self), as
    // self * n0 + nbr, sorted (the order ParTree::assemble creates them in).
    void edges(const std::vector<int>& cperm, const std::vector<int>& owner, int n0, std::vector<long long>& keys,
               hipStream_t s);
    // Every block in its edge's slot: dst[off[e] + (r - cstart[nbr]) + (c - cstart[self]) rows]
    // (the factor precision; dst zeroed here).
    template <typename S>
    void scatter_blocks(const std::vector<long long>& off, const std::vector<int>& cstart, const std::vector<int>& cols,
                        S* dst, size_t nel, hipStream_t s);
    // CSR of App (rowptr n + 1, col, val; hipMalloc'ed, the caller owns them).
    void csr(int** rowptr, int** col, double** val, hipStream_t s);
    // Distinct cluster pairs (s, p) joined by an entry of A^T A, as s * n0 + p, sorted.
    void ata_pairs(std::vector<long long>& keys, hipStream_t s);
    void release();

private:
    int n = 0, nnz = 0, n0 = 0;
    int *d_cp = nullptr, *d_ri = nullptr;     // A, CSC
    double* d_v = nullptr;
    int *d_rp = nullptr, *d_ci = nullptr;     // A, CSR
    int *d_gp = nullptr, *d_gi = nullptr;     // A^T A pattern, CSC
    long long gnnz = 0;
    int* d_pinv = nullptr;                    // permuted position of every natural index
    int* d_owner = nullptr;                   // level-0 cluster of every permuted index
    long long* d_ekeys = nullptr;             // sorted edge keys
    int ne = 0;
};

}  // namespace spaqr
