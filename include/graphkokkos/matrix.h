// graphkokkos/matrix.h -- drop-in for GraphKokkos's include/matrix.h: the CSR
// container spMtx<ValT>, the Matrix Market loader and the graph
// pre-processing helpers, on host memory, with no Kokkos dependency.
//
// Same member names and meaning as the reference (include/matrix.h:33-208):
// m, n, nz, matcode and the arrays Rst[m+1], Col[nz], Val[nz].  The arrays
// are gk::View<T>: reference-counted host buffers with Kokkos::View's
// shallow-copy semantics, element access by () and [], data(), size() and
// extent(0).  Products run on the GPU through libgraphhip (graphhip.h); the
// container itself never touches device memory.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "runtime.h"

#ifndef MM_TYPECODE_DEFINED
#define MM_TYPECODE_DEFINED
typedef char MM_typecode[4];
#endif

namespace gk {

template <class T>
class View {
public:
    View() = default;
    // zero-initialised, like View<T*>(label, n)
    View(const std::string &, size_t n) : n_(n), p_(n ? std::shared_ptr<T[]>(new T[n]()) : nullptr) {}
    T &operator()(size_t i) const { return p_[i]; }
    T &operator[](size_t i) const { return p_[i]; }
    T *data() const { return p_.get(); }
    size_t size() const { return n_; }
    size_t extent(int) const { return n_; }

private:
    size_t n_ = 0;
    std::shared_ptr<T[]> p_;
};

// Kokkos::resize: a new buffer holding the first min(old, n) elements, the
// rest zero.  Other views of the old buffer keep it.
template <class T>
void resize(View<T> &v, size_t n) {
    if (n == v.size()) return;
    View<T> w("", n);
    std::copy(v.data(), v.data() + std::min(n, v.size()), w.data());
    v = w;
}

// Kokkos::deep_copy between views of equal extent.
template <class T>
void deep_copy(const View<T> &dst, const View<T> &src) {
    if (dst.size() != src.size()) throw error(GK_EINVAL, "deep_copy: extents differ");
    std::copy(src.data(), src.data() + src.size(), dst.data());
}

}  // namespace gk

template <typename ValT>
class spMtx {
public:
    size_t m = 0;
    size_t n = 0;
    size_t nz = 0;
    MM_typecode matcode = {'M', 'C', 'I', 'G'};
    gk::View<int> Rst;
    gk::View<int> Col;
    gk::View<ValT> Val;

    // include/matrix.h:44-61, read by libgraphhip (gk_matrix_read): "mtx"
    // (1-based coordinates, self loops dropped, symmetric entries mirrored,
    // rows in file order, duplicates kept), "crs", "bin", "graph", "rmat"
    // (layouts and deviations in graphhip.h).  The reference's messages and
    // string throws are kept.
    spMtx(const char *filename, const std::string &format) {
        static const char *const names[][2] = {{"mtx", "MTX"}, {"crs", "CRS"}, {"bin", "BIN"},
                                               {"graph", "GRAPH"}, {"rmat", "RMAT"}};
        for (const auto &f : names)
            if (format == f[0]) {
                if (read_file(filename, f[0])) {
                    static thread_local std::string msg;
                    msg = std::string("Can't read ") + f[1] + " from file";
                    std::cout << msg << "\n";
                    throw msg.c_str();
                }
                return;
            }
    }

    spMtx() {}
    spMtx(size_t _m, size_t _n) : m(_m), n(_n), Rst("", _m + 1) {}
    spMtx(size_t _m, size_t _n, size_t _nz) : m(_m), n(_n), nz(_nz), Rst("", _m + 1), Col("", _nz), Val("", _nz) {}

    // copies share the arrays (include/matrix.h:72-75)
    spMtx(const spMtx &) = default;
    spMtx &operator=(const spMtx &) = default;
    spMtx(spMtx &&o) noexcept { *this = std::move(o); }
    spMtx &operator=(spMtx &&o) noexcept {
        if (this != &o) {
            m = o.m; n = o.n; nz = o.nz;
            std::memcpy(matcode, o.matcode, sizeof(MM_typecode));
            std::swap(Rst, o.Rst);
            std::swap(Col, o.Col);
            std::swap(Val, o.Val);
        }
        return *this;
    }

    void resizeRows(size_t newM) {
        m = newM;
        gk::resize(Rst, newM + 1);
    }
    void resizeVals(size_t newNz) {
        gk::resize(Col, newNz);
        gk::resize(Val, newNz);
        nz = newNz;
    }

    // pattern of src, values left zero (include/matrix.h:98-108)
    template <typename ValT2>
    void deep_copy_pattern(const spMtx<ValT2> &src) {
        m = src.m; n = src.n; nz = src.nz;
        Rst = gk::View<int>("", src.Rst.size());
        gk::deep_copy(Rst, src.Rst);
        Col = gk::View<int>("", src.Col.size());
        gk::deep_copy(Col, src.Col);
        Val = gk::View<ValT>("", src.nz);
    }

    void deep_copy(const spMtx &src) {
        m = src.m; n = src.n; nz = src.nz;
        std::memcpy(matcode, src.matcode, sizeof(MM_typecode));
        Rst = gk::View<int>("", src.Rst.size());
        Col = gk::View<int>("", src.Col.size());
        Val = gk::View<ValT>("", src.Val.size());
        gk::deep_copy(Rst, src.Rst);
        gk::deep_copy(Col, src.Col);
        gk::deep_copy(Val, src.Val);
    }

    // rows [begin, end) as a matrix of their own (include/matrix.h:151-163)
    spMtx extractRows(size_t begin, size_t end) const {
        const int b = Rst(begin), e = Rst(end);
        spMtx r(end - begin, n, size_t(e - b));
        for (size_t i = 0; i <= end - begin; ++i) r.Rst(i) = Rst(i + begin) - b;
        std::copy(Col.data() + b, Col.data() + e, r.Col.data());
        std::copy(Val.data() + b, Val.data() + e, r.Val.data());
        std::memcpy(r.matcode, matcode, sizeof(MM_typecode));
        return r;
    }

    bool operator==(const spMtx &o) const {
        if (m != o.m || n != o.n || nz != o.nz) return false;
        return std::equal(Rst.data(), Rst.data() + m + 1, o.Rst.data()) &&
               std::equal(Col.data(), Col.data() + nz, o.Col.data()) &&
               std::equal(Val.data(), Val.data() + nz, o.Val.data());
    }

    // 1-based "row col [value]" lines (include/matrix.h:177-189)
    void print_crs() const {
        std::cout << m << ' ' << nz << '\n';
        for (size_t i = 0; i < m; ++i)
            for (int j = Rst(i); j < Rst(i + 1); ++j) {
                std::cout << i + 1 << ' ' << Col(j) + 1;
                if (Val.size() > 0) std::cout << ' ' << Val(j);
                std::cout << '\n';
            }
    }

    // general coordinate Matrix Market file (include/matrix.h:225-252)
    int write_crs_to_mtx(const char *filename) const { return write_file(filename, "mtx", matcode[2] != 'P'); }

    // binary CRS (include/matrix.h:209-223), each array with its own data
    int write_crs_to_bin(const char *filename) const { return write_file(filename, "bin", true); }

    // Borrowed C-ABI view of the arrays; values only when ValT is int32 and
    // with_values is set (otherwise a pattern: every value 1).
    gk_csr csr(bool with_values = true) const {
        gk_csr c;
        c.m = (int64_t)m; c.n = (int64_t)n; c.nz = (int64_t)nz;
        c.rst = Rst.data();
        c.col = Col.data();
        c.val = nullptr;
        if constexpr (std::is_same<ValT, int>::value)
            if (with_values && Val.size() >= nz) c.val = Val.data();
        return c;
    }

private:
    int write_file(const char *filename, const char *format, bool with_values) const {
        std::vector<int32_t> v;
        if (with_values) {
            v.resize(nz);
            for (size_t j = 0; j < nz; ++j) v[j] = static_cast<int32_t>(Val(j));
        }
        const gk_csr c = csr(false);
        return gk_matrix_write(filename, format, &c, with_values ? v.data() : nullptr, matcode) == GK_OK ? 0 : -1;
    }

    int read_file(const char *filename, const char *format) {
        gk_mtx *mt = nullptr;
        if (gk_matrix_read(filename, format, &mt) != GK_OK) return -1;
        gk_csr p;
        const double *v = nullptr;
        char code[4];
        if (gk_mtx_view(mt, &p, &v, code) != GK_OK) { gk_mtx_free(mt); return -1; }
        m = (size_t)p.m; n = (size_t)p.n; nz = (size_t)p.nz;
        std::memcpy(matcode, code, 4);
        Rst = gk::View<int>("", m + 1);
        Col = gk::View<int>("", nz);
        Val = gk::View<ValT>("", nz);
        std::copy(p.rst, p.rst + m + 1, Rst.data());
        std::copy(p.col, p.col + nz, Col.data());
        for (size_t j = 0; j < nz; ++j) Val(j) = v ? static_cast<ValT>(v[j]) : ValT(1);
        gk_mtx_free(mt);
        return 0;
    }
};

// dst takes src's sizes, matcode and a copy of its arrays; resizing dst's
// own arrays first, so other spMtx sharing them keep the old contents
// (include/matrix.h:536-546).
template <typename ValT>
void deep_copy(spMtx<ValT> &dst, const spMtx<ValT> &src) {
    dst.m = src.m;
    dst.n = src.n;
    dst.resizeRows(src.m);
    dst.resizeVals(src.nz);
    std::copy(src.Rst.data(), src.Rst.data() + src.m + 1, dst.Rst.data());
    std::copy(src.Col.data(), src.Col.data() + src.nz, dst.Col.data());
    std::copy(src.Val.data(), src.Val.data() + src.nz, dst.Val.data());
    std::memcpy(dst.matcode, src.matcode, sizeof(MM_typecode));
}

// Dense row-major m x n matrix, Val(i * n + j), zero-initialised
// (include/matrix.h:548-630).  Copies share Val like the reference's
// Kokkos::View member; a sparse matrix assigned or converted to it is
// scattered into zeros.  One deliberate difference: copy-assignment between
// two distinct objects shares src's values here, while the reference's
// operator= returns before copying anything when the objects differ
// (include/matrix.h:575-577).
template <typename ValT>
class denseMtx {
public:
    size_t m = 0;
    size_t n = 0;
    gk::View<ValT> Val;

    denseMtx() {}
    denseMtx(size_t _m, size_t _n) : m(_m), n(_n), Val("", _m * _n) {}
    denseMtx(const denseMtx &src) : m(src.m), n(src.n), Val(src.Val) {}
    denseMtx(denseMtx &&mov) noexcept : m(mov.m), n(mov.n) { std::swap(Val, mov.Val); }
    template <typename ValT2>
    denseMtx(const spMtx<ValT2> &src) : m(src.m), n(src.n), Val("", src.m * src.n) {
        scatter(src);
    }

    denseMtx &operator=(const denseMtx &src) {
        if (this != &src) {
            m = src.m;
            n = src.n;
            Val = src.Val;
        }
        return *this;
    }
    denseMtx &operator=(denseMtx &&mov) noexcept {
        m = mov.m;
        n = mov.n;
        std::swap(Val, mov.Val);
        return *this;
    }
    // every entry reset to ValT(), then src's entries written in
    template <typename ValT2>
    denseMtx &operator=(const spMtx<ValT2> &src) {
        m = src.m;
        n = src.n;
        gk::resize(Val, m * n);
        std::fill(Val.data(), Val.data() + m * n, ValT());
        scatter(src);
        return *this;
    }

    bool operator==(const denseMtx &B) const {
        return m == B.m && n == B.n && Val.size() == B.Val.size() &&
               std::equal(Val.data(), Val.data() + Val.size(), B.Val.data());
    }

    // Kokkos::resize of the flat array: the first min(old, new) values kept
    void resize(size_t newM, size_t newN) {
        m = newM;
        n = newN;
        gk::resize(Val, newM * newN);
    }

    void print() const {
        for (size_t i = 0; i < m; ++i) {
            for (size_t j = 0; j < n; ++j) std::cout << std::setw(5) << Val(i * n + j) << ' ';
            std::cout << '\n';
        }
        std::cout << '\n';
    }

private:
    template <typename ValT2>
    void scatter(const spMtx<ValT2> &src) {
        for (size_t i = 0; i < src.m; ++i)
            for (int k = src.Rst(i); k < src.Rst(i + 1); ++k)
                Val(i * n + (size_t)src.Col(k)) = static_cast<ValT>(src.Val(k));
    }
};

// src's pattern with its values cast to ValT2; the matcode is not copied
// (include/matrix.h:632-641).
template <typename ValT2, typename ValT1>
spMtx<ValT2> convertType(const spMtx<ValT1> &src) {
    spMtx<ValT2> r;
    r.deep_copy_pattern(src);
    for (size_t j = 0; j < src.nz; ++j) r.Val(j) = static_cast<ValT2>(src.Val(j));
    return r;
}

// Counting-sort transpose; rows of the result come out sorted
// (include/matrix_la.h:155-179).
template <typename T>
spMtx<T> transpose(const spMtx<T> &A) {
    spMtx<T> R(A.n, A.m, A.nz);
    std::memcpy(R.matcode, A.matcode, sizeof(MM_typecode));
    for (size_t p = 0; p < A.nz; ++p) ++R.Rst(A.Col(p) + 1);
    for (size_t i = 0; i < A.n; ++i) R.Rst(i + 1) += R.Rst(i);
    std::vector<int> next(R.Rst.data(), R.Rst.data() + A.n);
    for (size_t i = 0; i < A.m; ++i)
        for (int p = A.Rst(i); p < A.Rst(i + 1); ++p) {
            const int q = next[A.Col(p)]++;
            R.Col(q) = (int)i;
            R.Val(q) = A.Val(p);
        }
    return R;
}

// Same pattern, every value 1, integer matcode (include/matrix.h:643-660).
template <typename T>
spMtx<int> build_adjacency_matrix(const spMtx<T> &G) {
    spMtx<int> R(G.m, G.n, G.nz);
    std::copy(G.Rst.data(), G.Rst.data() + G.m + 1, R.Rst.data());
    std::copy(G.Col.data(), G.Col.data() + G.nz, R.Col.data());
    std::fill(R.Val.data(), R.Val.data() + G.nz, 1);
    std::memcpy(R.matcode, G.matcode, sizeof(MM_typecode));
    R.matcode[2] = 'I';
    return R;
}

// Each row's leading run of columns below the diagonal; on sorted rows that
// is the strictly lower part (include/matrix.h:698-726).
template <typename T>
spMtx<T> extract_lower_triangle(const spMtx<T> &G) {
    spMtx<T> R(G.m, G.n);
    std::memcpy(R.matcode, G.matcode, sizeof(MM_typecode));
    for (size_t i = 0; i < G.m; ++i) {
        int p = G.Rst(i);
        while (p < G.Rst(i + 1) && (size_t)G.Col(p) < i) ++p;
        R.Rst(i + 1) = R.Rst(i) + (p - G.Rst(i));
    }
    R.nz = (size_t)R.Rst(G.m);
    R.Col = gk::View<int>("", R.nz);
    R.Val = gk::View<T>("", R.nz);
    for (size_t i = 0; i < G.m; ++i) {
        const int len = R.Rst(i + 1) - R.Rst(i);
        std::copy(G.Col.data() + G.Rst(i), G.Col.data() + G.Rst(i) + len, R.Col.data() + R.Rst(i));
        std::copy(G.Val.data() + G.Rst(i), G.Val.data() + G.Rst(i) + len, R.Val.data() + R.Rst(i));
    }
    return R;
}

// Row-wise merge of the sorted rows of L and L^T (include/matrix.h:662-696).
template <typename T>
spMtx<T> build_symm_from_lower(const spMtx<T> &Low) {
    spMtx<T> R(Low.m, Low.n, 2 * Low.nz);
    std::memcpy(R.matcode, Low.matcode, sizeof(MM_typecode));
    const spMtx<T> U = transpose(Low);
    size_t w = 0;
    for (size_t i = 0; i < Low.m; ++i) {
        int a = Low.Rst(i), b = U.Rst(i);
        const int ae = Low.Rst(i + 1), be = U.Rst(i + 1);
        while (a < ae || b < be) {
            const bool take_low = b >= be || (a < ae && Low.Col(a) < U.Col(b));
            if (take_low) { R.Col(w) = Low.Col(a); R.Val(w++) = Low.Val(a++); }
            else          { R.Col(w) = U.Col(b);   R.Val(w++) = U.Val(b++); }
        }
        R.Rst(i + 1) = (int)w;
    }
    return R;
}

// What remains of each row after its leading run of columns on or below
// the diagonal; on sorted rows that is the strictly upper part
// (include/matrix.h:728-755; the reference's version assigns raw arrays to
// its Views and does not compile when instantiated).
template <typename T>
spMtx<T> extract_upper_triangle(const spMtx<T> &G) {
    spMtx<T> R(G.m, G.n);
    std::memcpy(R.matcode, G.matcode, sizeof(MM_typecode));
    for (size_t i = 0; i < G.m; ++i) {
        int p = G.Rst(i);
        while (p < G.Rst(i + 1) && (size_t)G.Col(p) <= i) ++p;
        R.Rst(i + 1) = R.Rst(i) + (G.Rst(i + 1) - p);
    }
    R.nz = (size_t)R.Rst(G.m);
    R.Col = gk::View<int>("", R.nz);
    R.Val = gk::View<T>("", R.nz);
    for (size_t i = 0; i < G.m; ++i) {
        const int len = R.Rst(i + 1) - R.Rst(i);
        const int from = G.Rst(i + 1) - len;
        std::copy(G.Col.data() + from, G.Col.data() + from + len, R.Col.data() + R.Rst(i));
        std::copy(G.Val.data() + from, G.Val.data() + from + len, R.Val.data() + R.Rst(i));
    }
    return R;
}
