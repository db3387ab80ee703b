// Matrix Market files (the reference's include/mmio.hpp:174-377: sp_mmread, dense_mmread,
// vector_mmread, sp_mmwrite, dense_mmwrite) behind the C ABI.  Host code only.
//
// Read: coordinate files of type real, integer or pattern (pattern entries are 1) with property
// general, symmetric, hermitian (the mirror of every off-diagonal entry is added) or
// skew-symmetric (the mirror is added negated); duplicates are summed, as Eigen's
// setFromTriplets does; the result is CSC with rows sorted in each column.  Array files are
// read column-major (a vector is an n x 1 array).  Write: values with 17 significant digits,
// so a file read back gives the same doubles (the reference writes at the stream's default
// precision).
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/spaqr_gpu.h"
#include "errors.hpp"

namespace spaqr {
namespace {

struct MMError : std::runtime_error {
    int code;
    MMError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

std::string slurp(const char* path) {
    if (!path) throw MMError(-1, "no file name");
    FILE* f = std::fopen(path, "rb");
    if (!f) throw MMError(-1, std::string("cannot open ") + path + ": " + std::strerror(errno));
    std::string s;
    char buf[1 << 16];
    size_t r;
    while ((r = std::fread(buf, 1, sizeof(buf), f)) > 0) s.append(buf, r);
    std::fclose(f);
    return s;
}

std::string lower(std::string s) {
    for (auto& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

enum class Prop { general, symmetric, skew };

struct Parser {
    const std::string& s;
    size_t p = 0;
    explicit Parser(const std::string& str) : s(str) {}
    bool eof() const { return p >= s.size(); }
    // the next line that is neither empty nor a comment
    bool next_data_line(size_t& b, size_t& e) {
        while (p < s.size()) {
            b = p;
            e = s.find('\n', p);
            if (e == std::string::npos) e = s.size();
            p = e + 1;
            size_t q = b;
            while (q < e && (s[q] == ' ' || s[q] == '\t' || s[q] == '\r')) ++q;
            if (q < e && s[q] != '%') return true;
        }
        return false;
    }
};

struct Header {
    bool coordinate = true;
    bool pattern = false;
    Prop prop = Prop::general;
};

Header read_header(Parser& P) {
    size_t e = P.s.find('\n');
    if (e == std::string::npos) e = P.s.size();
    char tok[5][64] = {{0}};
    const std::string line = P.s.substr(0, e);
    if (std::sscanf(line.c_str(), "%63s %63s %63s %63s %63s", tok[0], tok[1], tok[2], tok[3], tok[4]) != 5 ||
        lower(tok[0]) != "%%matrixmarket" || lower(tok[1]) != "matrix")
        throw MMError(-2, "not a Matrix Market matrix header: " + line);
    Header h;
    const std::string fmt = lower(tok[2]), ty = lower(tok[3]), pr = lower(tok[4]);
    if (fmt == "coordinate") h.coordinate = true;
    else if (fmt == "array") h.coordinate = false;
    else throw MMError(-2, "unknown format " + fmt);
    if (ty == "pattern") h.pattern = true;
    else if (ty != "real" && ty != "integer" && ty != "double") throw MMError(-2, "unsupported type " + ty);
    if (pr == "general") h.prop = Prop::general;
    else if (pr == "symmetric" || pr == "hermitian") h.prop = Prop::symmetric;
    else if (pr == "skew-symmetric" || pr == "skew_symmetric") h.prop = Prop::skew;
    else throw MMError(-2, "unknown property " + pr);
    P.p = e + 1;
    return h;
}

struct Triplets {
    int nr = 0, nc = 0;
    std::vector<int> r, c;
    std::vector<double> v;
};

Triplets read_coordinate(const std::string& text) {
    Parser P(text);
    const Header h = read_header(P);
    if (!h.coordinate) throw MMError(-2, "expected a coordinate (sparse) file");
    size_t b, e;
    if (!P.next_data_line(b, e)) throw MMError(-3, "missing size line");
    long long M, N, K;
    if (std::sscanf(text.c_str() + b, "%lld %lld %lld", &M, &N, &K) != 3 || M < 0 || N < 0 || K < 0 ||
        M > 2147483647LL || N > 2147483646LL)
        throw MMError(-3, "malformed size line");
    Triplets t;
    t.nr = (int)M;
    t.nc = (int)N;
    t.r.reserve(h.prop == Prop::general ? K : 2 * K);
    long long got = 0;
    while (P.next_data_line(b, e)) {
        const char* q = text.c_str() + b;
        char* end = nullptr;
        const long long i = std::strtoll(q, &end, 10);
        if (end == q) throw MMError(-3, "malformed entry line " + std::to_string(got + 1));
        q = end;
        const long long j = std::strtoll(q, &end, 10);
        if (end == q) throw MMError(-3, "malformed entry line " + std::to_string(got + 1));
        q = end;
        double v = 1.0;
        if (!h.pattern) {
            v = std::strtod(q, &end);
            if (end == q) throw MMError(-3, "missing value on entry line " + std::to_string(got + 1));
        }
        if (i < 1 || i > M || j < 1 || j > N) throw MMError(-3, "entry index out of range on line " + std::to_string(got + 1));
        if (h.prop != Prop::general && i < j) throw MMError(-3, "entry above the diagonal in a symmetric file");
        if (h.prop == Prop::skew && i == j) throw MMError(-3, "diagonal entry in a skew-symmetric file");
        t.r.push_back((int)(i - 1));
        t.c.push_back((int)(j - 1));
        t.v.push_back(v);
        if (h.prop != Prop::general && i != j) {
            t.r.push_back((int)(j - 1));
            t.c.push_back((int)(i - 1));
            t.v.push_back(h.prop == Prop::skew ? -v : v);
        }
        ++got;
    }
    if (got != K) throw MMError(-4, "file holds " + std::to_string(got) + " entries, header says " + std::to_string(K));
    return t;
}

// CSC with rows sorted per column and duplicates summed
void to_csc(const Triplets& t, std::vector<int>& colptr, std::vector<int>& rowind, std::vector<double>& val) {
    const size_t m = t.r.size();
    std::vector<size_t> ord(m);
    for (size_t k = 0; k < m; ++k) ord[k] = k;
    std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) {
        return t.c[a] != t.c[b] ? t.c[a] < t.c[b] : t.r[a] < t.r[b];
    });
    colptr.assign((size_t)t.nc + 1, 0);
    rowind.clear();
    val.clear();
    for (size_t q = 0; q < m;) {
        const size_t k = ord[q];
        double s = 0.0;
        size_t q2 = q;
        while (q2 < m && t.c[ord[q2]] == t.c[k] && t.r[ord[q2]] == t.r[k]) s += t.v[ord[q2++]];
        rowind.push_back(t.r[k]);
        val.push_back(s);
        colptr[(size_t)t.c[k] + 1]++;
        q = q2;
    }
    for (int j = 0; j < t.nc; ++j) colptr[(size_t)j + 1] += colptr[j];
}

void read_array(const std::string& text, int& M, int& N, std::vector<double>* data) {
    Parser P(text);
    const Header h = read_header(P);
    if (h.coordinate) throw MMError(-2, "expected an array (dense) file");
    if (h.pattern) throw MMError(-2, "an array file cannot be of type pattern");
    if (h.prop != Prop::general) throw MMError(-2, "only general array files are read");
    size_t b, e;
    if (!P.next_data_line(b, e)) throw MMError(-3, "missing size line");
    long long m, n;
    if (std::sscanf(text.c_str() + b, "%lld %lld", &m, &n) != 2 || m < 0 || n < 0 || m * n > (1LL << 40))
        throw MMError(-3, "malformed size line");
    M = (int)m;
    N = (int)n;
    if (data) data->resize((size_t)m * n);
    long long got = 0;
    while (P.next_data_line(b, e)) {
        const char* q = text.c_str() + b;
        char* end = nullptr;
        const double v = std::strtod(q, &end);
        if (end == q) throw MMError(-3, "malformed value line " + std::to_string(got + 1));
        if (got < m * n && data) (*data)[(size_t)got] = v;
        ++got;
    }
    if (got != m * n) throw MMError(-4, "file holds " + std::to_string(got) + " values, header says " + std::to_string(m * n));
}

template <typename F>
int guarded(F&& f) {
    try {
        return f();
    } catch (MMError& e) {
        api_error() = e.what();
        return e.code;
    } catch (std::exception& e) {
        api_error() = e.what();
        return -100;
    }
}

}  // namespace
}  // namespace spaqr

using namespace spaqr;

extern "C" {

int spaqr_mm_read_sparse(const char* path, int* nrows, int* ncols, long long* nnz, int* colptr, int* rowind,
                         double* val) {
    return guarded([&] {
        const std::string text = slurp(path);
        const Triplets t = read_coordinate(text);
        std::vector<int> cp, ri;
        std::vector<double> v;
        to_csc(t, cp, ri, v);
        if (nrows) *nrows = t.nr;
        if (ncols) *ncols = t.nc;
        if (nnz) *nnz = (long long)ri.size();
        if (colptr) std::copy(cp.begin(), cp.end(), colptr);
        if (rowind) std::copy(ri.begin(), ri.end(), rowind);
        if (val) std::copy(v.begin(), v.end(), val);
        return 0;
    });
}

int spaqr_mm_read_dense(const char* path, int* nrows, int* ncols, double* data) {
    return guarded([&] {
        const std::string text = slurp(path);
        int M = 0, N = 0;
        std::vector<double> d;
        read_array(text, M, N, data ? &d : nullptr);
        if (nrows) *nrows = M;
        if (ncols) *ncols = N;
        if (data) std::copy(d.begin(), d.end(), data);
        return 0;
    });
}

int spaqr_mm_write_sparse(const char* path, int nrows, int ncols, const int* colptr, const int* rowind,
                          const double* val, int symmetric) {
    return guarded([&] {
        if (nrows < 0 || ncols < 0 || !colptr || (colptr[ncols] > 0 && (!rowind || !val)))
            throw MMError(-3, "bad matrix arguments");
        if (symmetric && nrows != ncols) throw MMError(-3, "a symmetric file needs a square matrix");
        long long cnt = 0;
        for (int j = 0; j < ncols; ++j)
            for (int k = colptr[j]; k < colptr[j + 1]; ++k) {
                if (rowind[k] < 0 || rowind[k] >= nrows) throw MMError(-3, "row index out of range");
                if (!symmetric || rowind[k] >= j) ++cnt;
            }
        FILE* f = path ? std::fopen(path, "wb") : nullptr;
        if (!f) throw MMError(-1, std::string("cannot open ") + (path ? path : "(null)") + " for writing");
        std::fprintf(f, "%%%%MatrixMarket matrix coordinate real %s\n", symmetric ? "symmetric" : "general");
        std::fprintf(f, "%d %d %lld\n", nrows, ncols, cnt);
        for (int j = 0; j < ncols; ++j)
            for (int k = colptr[j]; k < colptr[j + 1]; ++k)
                if (!symmetric || rowind[k] >= j) std::fprintf(f, "%d %d %.17g\n", rowind[k] + 1, j + 1, val[k]);
        const bool ok = std::fclose(f) == 0;
        if (!ok) throw MMError(-1, "write failed");
        return 0;
    });
}

int spaqr_mm_write_dense(const char* path, int nrows, int ncols, const double* data) {
    return guarded([&] {
        if (nrows < 0 || ncols < 0 || (!data && (size_t)nrows * ncols > 0)) throw MMError(-3, "bad matrix arguments");
        FILE* f = path ? std::fopen(path, "wb") : nullptr;
        if (!f) throw MMError(-1, std::string("cannot open ") + (path ? path : "(null)") + " for writing");
        std::fprintf(f, "%%%%MatrixMarket matrix array real general\n%d %d\n", nrows, ncols);
        for (size_t e = 0; e < (size_t)nrows * ncols; ++e) std::fprintf(f, "%.17g\n", data[e]);
        const bool ok = std::fclose(f) == 0;
        if (!ok) throw MMError(-1, "write failed");
        return 0;
    });
}

}  // extern "C"
