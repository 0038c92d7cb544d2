// CSR construction, host SpMV, model validation and fingerprint.
#include "lp.hpp"
#include "parallel.hpp"

#include <algorithm>
#include <cstdio>
#include <stdexcept>
#include <unordered_set>

namespace nyk {

// src/model.cpp:11-38.  Counting sort by row, then a stable sort of each row
// by column, so duplicate entries are summed in their input order.
Csr Csr::from_entries(int64_t rows, int64_t cols, std::vector<Entry> e) {
    if (rows < 0 || cols < 0) throw std::runtime_error("Negative matrix dimension");
    for (const auto &a : e)
        if (a.row < 0 || a.row >= rows || a.col < 0 || a.col >= cols || !std::isfinite(a.value))
            throw std::runtime_error("Invalid sparse entry");
    std::vector<int64_t> start(rows + 1, 0);
    for (const auto &a : e) start[a.row + 1]++;
    for (int64_t i = 0; i < rows; ++i) start[i + 1] += start[i];
    std::vector<Entry> sorted(e.size());
    {
        std::vector<int64_t> pos(start.begin(), start.end() - 1);
        for (const auto &a : e) sorted[pos[a.row]++] = a;
    }
    Csr s;
    s.rows = rows;
    s.cols = cols;
    s.ptr.assign(rows + 1, 0);
    s.index.reserve(e.size());
    s.value.reserve(e.size());
    for (int64_t i = 0; i < rows; ++i) {
        auto first = sorted.begin() + start[i], last = sorted.begin() + start[i + 1];
        std::stable_sort(first, last, [](const Entry &a, const Entry &b) { return a.col < b.col; });
        for (auto it = first; it != last;) {
            int64_t j = it->col;
            double v = it->value;
            for (++it; it != last && it->col == j; ++it) v += it->value;
            if (!std::isfinite(v)) throw std::runtime_error("Coefficient sum overflow");
            if (v != 0) {
                s.index.push_back(j);
                s.value.push_back(v);
            }
        }
        s.ptr[i + 1] = int64_t(s.index.size());
    }
    return s;
}

// src/model.cpp:39-58 (entries of each transposed row keep ascending row order).
Csr Csr::transposed() const {
    Csr t;
    t.rows = cols;
    t.cols = rows;
    t.ptr.assign(cols + 1, 0);
    t.index.resize(value.size());
    t.value.resize(value.size());
    for (auto j : index) t.ptr[j + 1]++;
    for (int64_t j = 0; j < cols; ++j) t.ptr[j + 1] += t.ptr[j];
    std::vector<int64_t> pos(t.ptr.begin(), t.ptr.end() - 1);
    for (int64_t i = 0; i < rows; ++i)
        for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k) {
            int64_t p = pos[index[k]]++;
            t.index[p] = i;
            t.value[p] = value[k];
        }
    return t;
}

// src/model.cpp:59-73.  Each row is summed left to right by one thread, so the
// result does not depend on the thread count.
void Csr::multiply(const double *x, double *y) const {
    parallel_for(rows, rows > 10000, [&](int64_t i) {
        double v = 0;
        for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k) v += value[k] * x[index[k]];
        y[i] = v;
    });
}

// src/model.cpp:74-109 (continuous LP subset).
void Model::validate() const {
    const size_t n = c.size(), m = rl.size();
    if (lb.size() != n || ub.size() != n || ru.size() != m || A.rows != int64_t(m) || A.cols != int64_t(n) ||
        (!names.empty() && names.size() != n) || (!row_names.empty() && row_names.size() != m))
        throw std::runtime_error("Inconsistent model dimensions");
    if (A.ptr.size() != m + 1 || A.ptr.front() != 0 || A.ptr.back() != A.nnz() || A.index.size() != A.value.size())
        throw std::runtime_error("Invalid CSR storage");
    if (!std::isfinite(offset) || (sense != 1 && sense != -1)) throw std::runtime_error("Invalid objective offset/sense");
    for (size_t j = 0; j < n; ++j) {
        if (!std::isfinite(c[j])) throw std::runtime_error("Nonfinite objective coefficient");
        if (std::isnan(lb[j]) || std::isnan(ub[j]) || lb[j] == kInf || ub[j] == -kInf || lb[j] > ub[j])
            throw std::runtime_error("Invalid bounds for " + column_name(int64_t(j)));
    }
    for (size_t i = 0; i < m; ++i) {
        if (std::isnan(rl[i]) || std::isnan(ru[i]) || rl[i] == kInf || ru[i] == -kInf || rl[i] > ru[i])
            throw std::runtime_error("Invalid row bounds");
        if (A.ptr[i] > A.ptr[i + 1] || A.ptr[i] < 0 || A.ptr[i + 1] > A.nnz())
            throw std::runtime_error("Invalid CSR offsets");
    }
    for (int64_t k = 0; k < A.nnz(); ++k)
        if (A.index[k] < 0 || A.index[k] >= int64_t(n) || !std::isfinite(A.value[k]))
            throw std::runtime_error("Invalid matrix coefficient/index");
    for (const auto *list : {&names, &row_names}) {
        std::unordered_set<std::string> seen;
        seen.reserve(list->size());
        for (const auto &s : *list)
            if (s.empty() || !seen.insert(s).second) throw std::runtime_error("Duplicate/empty variable or row name");
    }
}

// src/model.cpp:274-309: FNV-1a 64 over the raw bytes of dims, offset, sense,
// c, q (all zero for an LP), lb, ub, rl, ru, A.value, A.ptr, A.index and the
// variable types (all continuous), printed as lowercase hex.
std::string Model::fingerprint() const {
    uint64_t h = 14695981039346656037ULL;
    auto bytes = [&](const void *data, size_t size) {
        auto *p = static_cast<const unsigned char *>(data);
        for (size_t i = 0; i < size; ++i) {
            h ^= p[i];
            h *= 1099511628211ULL;
        }
    };
    auto span = [&](const auto &v) {
        if (!v.empty()) bytes(v.data(), v.size() * sizeof(v[0]));
    };
    bytes(&A.rows, sizeof A.rows);
    bytes(&A.cols, sizeof A.cols);
    bytes(&offset, sizeof offset);
    bytes(&sense, sizeof sense);
    span(c);
    const double zero = 0;
    for (size_t j = 0; j < c.size(); ++j) bytes(&zero, sizeof zero);
    span(lb);
    span(ub);
    span(rl);
    span(ru);
    span(A.value);
    span(A.ptr);
    span(A.index);
    const int continuous = 0;
    for (size_t j = 0; j < c.size(); ++j) bytes(&continuous, sizeof continuous);
    char text[24];
    std::snprintf(text, sizeof text, "%llx", (unsigned long long)h);
    return text;
}

}  // namespace nyk
