// sparse_io.cpp -- host-side sparse matrix inputs for the CG path: Matrix
// Market files, the COO -> CSR conversion, the reference's random symmetric
// tridiagonal test matrix, and CSR -> DIA.
//
// Follows the reference's readers (conjugateGradient/matrix/sparse_matrix.h:
// CooMatrix::InitMarket :217-390, CsrMatrix::Init :666-734) and generator
// (conjugateGradient/matrix/tridiag.cpp:11-47): same entry order (entries
// stable-sorted by row, then column; duplicates kept), same treatment of
// "symmetric"/"skew" (the mirrored entry follows its original, negated for
// skew), "pattern" entries take the default value 1, and "array" files are
// read column-major.  Errors are returned, never exit().
#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "perks_amd.h"

namespace perks {
namespace {

struct Coo {
    int32_t row, col;
    double val;
};

// Reads one line (any length) into `line`; false at end of file.
bool read_line(FILE* f, std::vector<char>& line)
{
    line.clear();
    int c;
    while ((c = std::fgetc(f)) != EOF) {
        if (c == '\n') break;
        line.push_back((char)c);
    }
    if (c == EOF && line.empty()) return false;
    line.push_back('\0');
    return true;
}

int read_market(const char* path, int32_t* nrows, int32_t* ncols, std::vector<Coo>* out)
{
    FILE* f = std::fopen(path, "r");
    if (!f) return PERKS_ERR_INVALID_ARG;
    bool array = false, symmetric = false, skew = false, have_size = false;
    long long declared = 0;
    std::vector<char> line;
    std::vector<Coo>& coo = *out;
    coo.clear();
    int rc = PERKS_OK;
    long long rows = 0, cols = 0;
    while (read_line(f, line)) {
        const char* s = line.data();
        if (s[0] == '%') {
            if (s[1] == '%') {
                symmetric = std::strstr(s, "symmetric") != nullptr;
                skew = std::strstr(s, "skew") != nullptr;
                array = std::strstr(s, "array") != nullptr;
            }
            continue;
        }
        if (!have_size) {
            long long a = 0, b = 0, c = 0;
            const int np = std::sscanf(s, "%lld %lld %lld", &a, &b, &c);
            if (np <= 0) continue;  // blank line before the size line
            if (!array && np == 3) {
                declared = symmetric ? 2 * c : c;
            } else if (array && np >= 2) {
                declared = a * b;
            } else {
                rc = PERKS_ERR_INVALID_ARG;
                break;
            }
            if (a <= 0 || b <= 0 || a > 0x7fffffffLL || b > 0x7fffffffLL || declared < 0) {
                rc = PERKS_ERR_UNSUPPORTED_SIZE;
                break;
            }
            rows = a;
            cols = b;
            coo.reserve((size_t)std::min<long long>(declared, 1LL << 31));
            have_size = true;
            continue;
        }
        if ((long long)coo.size() >= declared) {
            rc = PERKS_ERR_INVALID_ARG;  // more entries than declared
            break;
        }
        if (array) {
            char* end = nullptr;
            const double v = std::strtod(s, &end);
            if (end == s) continue;  // blank line
            const long long k = (long long)coo.size();
            coo.push_back(Coo{(int32_t)(k % rows), (int32_t)(k / rows), v});
            continue;
        }
        char* t = nullptr;
        const long long r = std::strtol(s, &t, 0);
        if (t == s) continue;  // blank line
        const char* l = t;
        const long long c = std::strtol(l, &t, 0);
        if (t == l) {
            rc = PERKS_ERR_INVALID_ARG;
            break;
        }
        l = t;
        double v = std::strtod(l, &t);
        if (t == l) v = 1.0;  // pattern entry: the default value
        if (r < 1 || c < 1 || r > rows || c > cols) {
            rc = PERKS_ERR_INVALID_ARG;
            break;
        }
        coo.push_back(Coo{(int32_t)(r - 1), (int32_t)(c - 1), v});
        if (symmetric && r != c) coo.push_back(Coo{(int32_t)(c - 1), (int32_t)(r - 1), skew ? -v : v});
    }
    std::fclose(f);
    if (rc == PERKS_OK && !have_size) rc = PERKS_ERR_INVALID_ARG;
    if (rc != PERKS_OK) return rc;
    *nrows = (int32_t)rows;
    *ncols = (int32_t)cols;
    return PERKS_OK;
}

// Stable sort by (row, col), then row offsets: the reference's CsrMatrix::Init.
void coo_to_csr(std::vector<Coo>& coo, int32_t n, int32_t* rowptr, int32_t* col, double* val)
{
    std::stable_sort(coo.begin(), coo.end(), [](const Coo& a, const Coo& b) {
        return a.row < b.row || (a.row == b.row && a.col < b.col);
    });
    int64_t k = 0;
    for (int32_t r = 0; r <= n; ++r) {
        while (k < (int64_t)coo.size() && coo[k].row < r) ++k;
        rowptr[r] = (int32_t)k;
    }
    for (size_t i = 0; i < coo.size(); ++i) {
        col[i] = coo[i].col;
        val[i] = coo[i].val;
    }
}

// The reference draws from the C library's rand() in a fresh process (seed
// 1).  The same sequence comes from glibc's random_r on a private state
// seeded with 1, which leaves the caller's rand() state alone.
struct RandSeq {
    random_data rd;
    char state[128];
    RandSeq()
    {
        std::memset(&rd, 0, sizeof(rd));
        std::memset(state, 0, sizeof(state));
        initstate_r(1u, state, sizeof(state), &rd);
    }
    int next()
    {
        int32_t r = 0;
        random_r(&rd, &r);
        return (int)r;
    }
};

template <typename T>
void gen_tridiag(int32_t n, int32_t* I, int32_t* J, T* val)
{
    // rand() / RAND_MAX with T operands: RAND_MAX converts to T, as the
    // reference's implicit conversion does (2^31 - 1 rounds to 2^31 in float).
    RandSeq rs;
    const int32_t nz = (n - 2) * 3 + 4;
    I[0] = 0;
    J[0] = 0;
    J[1] = 1;
    val[0] = static_cast<T>(rs.next()) / static_cast<T>(RAND_MAX) + 10.0f;
    val[1] = static_cast<T>(rs.next()) / static_cast<T>(RAND_MAX);
    for (int32_t i = 1; i < n; ++i) {
        I[i] = i > 1 ? I[i - 1] + 3 : 2;
        const int32_t start = (i - 1) * 3 + 2;
        J[start] = i - 1;
        J[start + 1] = i;
        if (i < n - 1) J[start + 2] = i + 1;
        val[start] = val[start - 1];
        val[start + 1] = static_cast<T>(rs.next()) / static_cast<T>(RAND_MAX) + 10.0f;
        if (i < n - 1) val[start + 2] = static_cast<T>(rs.next()) / static_cast<T>(RAND_MAX);
    }
    I[n] = nz;
}

}  // namespace
}  // namespace perks

extern "C" {

int perks_mm_read_csr(const char* path, int32_t* n, int64_t* nnz, int32_t** rowptr, int32_t** col, double** val)
{
    if (!path || !n || !nnz || !rowptr || !col || !val) return PERKS_ERR_INVALID_ARG;
    std::vector<perks::Coo> coo;
    int32_t rows = 0, cols = 0;
    int rc = perks::read_market(path, &rows, &cols, &coo);
    if (rc != PERKS_OK) return rc;
    if (coo.size() > 0x7fffffffULL) return PERKS_ERR_UNSUPPORTED_SIZE;
    int32_t* rp = (int32_t*)std::malloc(sizeof(int32_t) * ((size_t)rows + 1));
    int32_t* cl = (int32_t*)std::malloc(sizeof(int32_t) * std::max<size_t>(coo.size(), 1));
    double* vl = (double*)std::malloc(sizeof(double) * std::max<size_t>(coo.size(), 1));
    if (!rp || !cl || !vl) {
        std::free(rp);
        std::free(cl);
        std::free(vl);
        return PERKS_ERR_INVALID_ARG;
    }
    perks::coo_to_csr(coo, rows, rp, cl, vl);
    *n = rows;
    *nnz = (int64_t)coo.size();
    *rowptr = rp;
    *col = cl;
    *val = vl;
    (void)cols;
    return PERKS_OK;
}

void perks_free_host(void* p) { std::free(p); }

int perks_gen_tridiag_f64(int32_t n, int32_t* rowptr, int32_t* col, double* val)
{
    if (n < 2 || !rowptr || !col || !val) return PERKS_ERR_INVALID_ARG;
    perks::gen_tridiag<double>(n, rowptr, col, val);
    return PERKS_OK;
}

int perks_gen_tridiag_f32(int32_t n, int32_t* rowptr, int32_t* col, float* val)
{
    if (n < 2 || !rowptr || !col || !val) return PERKS_ERR_INVALID_ARG;
    perks::gen_tridiag<float>(n, rowptr, col, val);
    return PERKS_OK;
}

int perks_csr_to_dia(int32_t n, const int32_t* rowptr, const int32_t* col, const double* val, int32_t max_diag,
                     int32_t* offsets, double* dia_val)
{
    if (n <= 0 || !rowptr || !col || !val || !offsets || max_diag <= 0) return PERKS_ERR_INVALID_ARG;
    // Distinct offsets col - row, ascending; each entry goes to its diagonal.
    std::vector<int64_t> offs;
    for (int32_t r = 0; r < n; ++r)
        for (int32_t k = rowptr[r]; k < rowptr[r + 1]; ++k) offs.push_back((int64_t)col[k] - r);
    std::sort(offs.begin(), offs.end());
    offs.erase(std::unique(offs.begin(), offs.end()), offs.end());
    if ((int32_t)offs.size() > max_diag) return PERKS_ERR_UNSUPPORTED;
    const int32_t nd = (int32_t)offs.size();
    for (int32_t d = 0; d < nd; ++d) offsets[d] = (int32_t)offs[d];
    if (dia_val) {
        std::memset(dia_val, 0, sizeof(double) * (size_t)nd * n);
        for (int32_t r = 0; r < n; ++r)
            for (int32_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
                const int64_t o = (int64_t)col[k] - r;
                const int32_t d = (int32_t)(std::lower_bound(offs.begin(), offs.end(), o) - offs.begin());
                dia_val[(size_t)d * n + r] += val[k];  // duplicates of one entry add up, as in a product
            }
    }
    return nd;
}

}  // extern "C"
