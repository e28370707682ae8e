// pqb_oracle.cpp -- CPU oracle for the permanent engine.  TEST INFRASTRUCTURE:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load it, and only as the checker (or the timed CPU baseline), never as the
// product path.
//
// Part 1 restates the reference algorithms step by step, in the reference's
// arithmetic order, for the scalar types the reference instantiates (double,
// long double) plus double-double and binary128 for extended-precision
// checks:
//   bbfg_plain      BBFGPermanentCalculator.hpp:87-235 (+ .cpp:48-59, 88-126)
//   bbfg_repeated   BBFGPermanentCalculatorRepeated.hpp:74-307 (+ .cpp:74-141)
//   nary_gray       common/source/n_aryGrayCodeCounter.cpp:44-248
//   binom128        common/source/common_functionalities.cpp:131-160
//   glynn_recursive GlynnPermanentCalculator.hpp:47-201
//   glynn_repeated  GlynnPermanentCalculatorRepeated.hpp:49-258
//   glynn_simple    GlynnPermanentCalculatorSimple.hpp:61-168
//   complexm        common/source/PicComplexM.hpp:29-201 (order buckets)
// The reference runs the chunks of bbfg_plain / bbfg_repeated on TBB threads
// and combines thread-local sums in an unspecified order; here every chunk has
// its own partial and partials are combined in chunk order, so the result
// does not depend on the number of worker threads.
//
// Part 2 holds independent checks that share no code path with part 1:
// permutation-sum brute force, Ryser's formula (Gray-code subsets, any
// scalar type) and an exact signed-int128 Ryser for integer matrices.
//
// Matrices are row-major complex128 with a row stride `ld` (in complex
// elements).  Results are returned as (re, im) doubles, and for the
// double-double / binary128 types additionally as (hi, lo) pairs.

#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <vector>

namespace {

// ----------------------------------------------------------- scalar types --

struct DD {  // double-double, value = hi + lo
    double hi = 0.0, lo = 0.0;
    DD() = default;
    DD(double h) : hi(h), lo(0.0) {}
    DD(double h, double l) : hi(h), lo(l) {}
};

inline DD two_sum(double a, double b) {
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    return DD(s, e);
}
inline DD fast_two_sum(double a, double b) {
    double s = a + b;
    return DD(s, b - (s - a));
}
inline DD operator+(DD a, DD b) {
    DD s = two_sum(a.hi, b.hi);
    DD t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return fast_two_sum(s.hi, s.lo);
}
inline DD operator-(DD a) { return DD(-a.hi, -a.lo); }
inline DD operator-(DD a, DD b) { return a + (-b); }
inline DD operator*(DD a, DD b) {
    double p = a.hi * b.hi;
    double e = std::fma(a.hi, b.hi, -p);
    e += a.hi * b.lo + a.lo * b.hi;
    return fast_two_sum(p, e);
}
inline DD operator/(DD a, DD b) {  // only used for exact powers of two here
    double q1 = a.hi / b.hi;
    DD r = a - DD(q1) * b;
    double q2 = r.hi / b.hi;
    return fast_two_sum(q1, q2);
}
inline DD& operator+=(DD& a, DD b) { return a = a + b; }
inline DD& operator-=(DD& a, DD b) { return a = a - b; }
inline DD& operator*=(DD& a, DD b) { return a = a * b; }
inline bool operator>(DD a, DD b) { return a.hi > b.hi || (a.hi == b.hi && a.lo > b.lo); }
inline bool operator<(DD a, DD b) { return b > a; }

typedef __float128 Q;

template <class T> inline double hi_of(T v) { return (double)v; }
template <class T> inline double lo_of(T v) { return (double)(v - (T)(double)v); }
template <> inline double hi_of<DD>(DD v) { return v.hi + v.lo; }
template <> inline double lo_of<DD>(DD v) { return v.lo - ((v.hi + v.lo) - v.hi); }

template <class T> inline T from_d(double v) { return T(v); }

// ------------------------------------------------------------------ complex --

// Complex arithmetic in the order of Complex_base (PicComplex.hpp:143-158,
// 201-258): (a+ib)(c+id) = (ac - bd) + i(ad + bc).
template <class T>
struct Cx {
    T re, im;
    Cx() : re(0.0), im(0.0) {}
    Cx(T r, T i) : re(r), im(i) {}
};
template <class T> inline Cx<T> operator*(const Cx<T>& a, const Cx<T>& b) {
    return Cx<T>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <class T> inline Cx<T>& operator*=(Cx<T>& a, const Cx<T>& b) {
    T r = a.re * b.re - a.im * b.im;
    a.im = a.re * b.im + a.im * b.re;
    a.re = r;
    return a;
}
template <class T> inline Cx<T>& operator+=(Cx<T>& a, const Cx<T>& b) {
    a.re = a.re + b.re;
    a.im = a.im + b.im;
    return a;
}
template <class T> inline Cx<T>& operator-=(Cx<T>& a, const Cx<T>& b) {
    a.re = a.re - b.re;
    a.im = a.im - b.im;
    return a;
}
template <class T> inline Cx<T> operator+(Cx<T> a, const Cx<T>& b) { return a += b; }
template <class T> inline Cx<T> operator-(Cx<T> a, const Cx<T>& b) { return a -= b; }
template <class T> inline Cx<T> scale(const Cx<T>& a, T s) { return Cx<T>(a.re * s, a.im * s); }
template <class T> inline T mag2(const Cx<T>& a) { return a.re * a.re + a.im * a.im; }

template <class T>
std::vector<Cx<T>> load(const double* A, int R, int C, int64_t ld) {
    std::vector<Cx<T>> m((size_t)R * C);
    for (int i = 0; i < R; ++i)
        for (int j = 0; j < C; ++j)
            m[(size_t)i * C + j] = Cx<T>(from_d<T>(A[2 * (i * ld + j)]), from_d<T>(A[2 * (i * ld + j) + 1]));
    return m;
}

template <class T>
void store(const Cx<T>& v, double* out) {
    out[0] = hi_of(v.re);
    out[1] = hi_of(v.im);
    out[2] = hi_of(v.re);
    out[3] = lo_of(v.re);
    out[4] = hi_of(v.im);
    out[5] = lo_of(v.im);
}

// ------------------------------------------------------------ thread pool --

void par_for(int64_t njobs, int nthreads, const std::function<void(int64_t)>& fn) {
    if (nthreads <= 1 || njobs <= 1) {
        for (int64_t j = 0; j < njobs; ++j) fn(j);
        return;
    }
    std::atomic<int64_t> next(0);
    std::vector<std::thread> th;
    const int nt = (int)std::min<int64_t>(nthreads, njobs);
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&]() {
            for (;;) {
                int64_t j = next.fetch_add(1);
                if (j >= njobs) break;
                fn(j);
            }
        });
    for (auto& t : th) t.join();
}

// --------------------------------------------------------------- ComplexM --

// Order-bucketed accumulator, PicComplexM.hpp:29-201 (ORDER_CUTOFF = 1e-5).
template <class T>
struct ComplexM {
    std::vector<Cx<T>> addends;
    static bool gt(const Cx<T>& a, const Cx<T>& b) { return mag2(a) > mag2(b); }  // PicComplex.hpp:298-306
    static bool lt(const Cx<T>& a, const Cx<T>& b) { return mag2(a) < mag2(b); }
    void add(const Cx<T>& value, size_t index) {
        if (addends.size() <= index) {
            addends.push_back(value);
            return;
        }
        Cx<T> tmp = addends[index];
        if (gt(scale(tmp, T(1e-5)), value)) {
            addends[index] = value;
            add(tmp, index + 1);
        } else if (lt(tmp, scale(value, T(1e-5)))) {
            add(value, index + 1);
        } else {
            addends[index] += value;
        }
    }
    void subtract(const Cx<T>& value, size_t index) {
        Cx<T> neg(T(0.0) - value.re, T(0.0) - value.im);
        if (addends.size() <= index) {
            addends.push_back(neg);
            return;
        }
        Cx<T> tmp = addends[index];
        if (gt(scale(tmp, T(1e-5)), value)) {
            addends[index] = neg;
            add(tmp, index + 1);
        } else if (lt(tmp, scale(value, T(1e-5)))) {
            subtract(value, index + 1);
        } else {
            addends[index] -= value;
        }
    }
    void operator+=(const Cx<T>& v) { add(v, 0); }
    Cx<T> get() const {
        Cx<T> s;
        for (const auto& a : addends) s += a;
        return s;
    }
};

// ------------------------------------------------------------ BBFG plain --

// BBFGPermanentCalculator_Tasks<matrix_type, scalar_type, precision>::calculate
// (BBFGPermanentCalculator.hpp:87-235); `njobs` plays the reference's
// `concurrency` (= 4 * hardware threads there, :120-122).
template <class T>
Cx<T> bbfg_plain(const double* Ad, int R, int C, int64_t ld, int64_t njobs, int nthreads,
                 int64_t job_lo = 0, int64_t job_hi = -1) {
    if (R == 0 || C == 0) return Cx<T>(T(1.0), T(0.0));  // :90-93
    if (R >= C + 2) return Cx<T>(T(0.0), T(0.0));        // :94-95
    std::vector<Cx<T>> mtx = load<T>(Ad, R, C, ld);
    std::vector<Cx<T>> mtx2(mtx.size());
    for (size_t i = 0; i < mtx.size(); ++i) mtx2[i] = scale(mtx[i], T(2.0));  // :107-110
    const uint64_t Idx_max = 1ULL << (R - 1);
    int64_t concurrency = std::min<int64_t>(njobs, (int64_t)Idx_max);
    if (concurrency < 1) concurrency = 1;
    if (job_hi < 0 || job_hi > concurrency) job_hi = concurrency;
    std::vector<Cx<T>> part((size_t)concurrency);
    par_for(job_hi - job_lo, nthreads, [&](int64_t jj) {
        const int64_t job_idx = job_lo + jj;
        const int64_t work_batch = Idx_max / concurrency;
        const int64_t initial_offset = job_idx * work_batch;
        int64_t offset_max = (job_idx + 1) * work_batch;
        if (job_idx == concurrency - 1) offset_max = Idx_max;
        std::vector<int> delta_vec(R, 0);
        int parity = 1;
        std::vector<Cx<T>> colsum(mtx.begin(), mtx.begin() + C);  // row 0
        uint64_t g = initial_offset ^ (initial_offset >> 1);
        for (int row_idx = R - 1; row_idx > 0; --row_idx) {  // :150-167
            if (g & 1) {
                for (int c = 0; c < C; ++c) colsum[c] -= mtx[(size_t)row_idx * C + c];
                parity = -parity;
                delta_vec[row_idx] = 1;
            } else {
                for (int c = 0; c < C; ++c) colsum[c] += mtx[(size_t)row_idx * C + c];
                delta_vec[row_idx] = 0;
            }
            g >>= 1;
        }
        Cx<T> addend;
        {
            Cx<T> pr = colsum[0];  // product_reduction, BBFGPermanentCalculator.cpp:48-59
            for (int c = 1; c < C; ++c) pr *= colsum[c];
            if (parity == 1) addend += pr; else addend -= pr;  // :171-176
        }
        for (int64_t idx = initial_offset + 1; idx < offset_max; ++idx) {  // :179-221
            uint64_t idx_loc = (uint64_t)idx;
            int change_idx = R - 1;
            while ((idx_loc & 1ULL) == 0) {
                idx_loc >>= 1;
                --change_idx;
            }
            int& d = delta_vec[change_idx];
            d = d ? 0 : 1;
            const Cx<T>* row2 = &mtx2[(size_t)change_idx * C];
            parity = -parity;
            Cx<T> prod(T((double)parity), T(0.0));
            if (d) {
                for (int c = 0; c < C; ++c) {
                    colsum[c] -= row2[c];
                    prod *= colsum[c];
                }
            } else {
                for (int c = 0; c < C; ++c) {
                    colsum[c] += row2[c];
                    prod *= colsum[c];
                }
            }
            addend += prod;
        }
        part[(size_t)job_idx] = addend;
    });
    Cx<T> perm;
    for (int64_t j = job_lo; j < job_hi; ++j) perm += part[(size_t)j];
    const T div = T((double)(1ULL << (R - 1)));  // :232
    perm.re = perm.re / div;
    perm.im = perm.im / div;
    return perm;
}

// --------------------------------------------------- n-ary Gray counter --

// n_aryGrayCodeCounter.cpp:44-248.
struct NaryGray {
    std::vector<int> limits, chain, gray;
    int64_t offset = 0, offset_max = 0;
    bool bad = false;
    NaryGray(const std::vector<int>& lim, int64_t initial_offset) : limits(lim) {
        if (limits.empty()) {
            offset_max = 0;
            offset = 0;
            return;
        }
        offset_max = limits[0];
        for (size_t i = 1; i < limits.size(); ++i) offset_max *= limits[i];
        offset_max--;
        offset = initial_offset;
        init(initial_offset);
    }
    void init(int64_t off) {  // :115-140
        if (off < 0 || off > offset_max) {
            bad = true;  // reference throws "Wrong value of initial_offset"
            return;
        }
        chain.assign(limits.size(), 0);
        for (size_t i = 0; i < limits.size(); ++i) {
            chain[i] = (int)(off % limits[i]);
            off /= limits[i];
        }
        gray.assign(limits.size(), 0);
        int parity = 0;
        for (int64_t j = (int64_t)limits.size() - 1; j >= 0; --j) {
            gray[j] = parity ? limits[j] - 1 - chain[j] : chain[j];
            parity ^= (gray[j] & 1);
        }
    }
    // returns 1 when exhausted (:198-200)
    int next(int& changed, int& prev, int& val) {
        changed = 0;
        if (offset >= offset_max) return 1;
        bool upd = true;
        size_t ci = 0;
        while (upd) {
            if (chain[ci] < limits[ci] - 1) {
                chain[ci]++;
                upd = false;
            } else if (chain[ci] == limits[ci] - 1) {
                chain[ci] = 0;
                upd = true;
            }
            ci++;
        }
        int parity = 0;
        for (int64_t j = (int64_t)limits.size() - 1; j >= 0; --j) {
            int nv = parity ? limits[j] - 1 - chain[j] : chain[j];
            parity ^= (nv & 1);
            if (nv != gray[j]) {
                prev = gray[j];
                val = nv;
                gray[j] = nv;
                changed = (int)j;
                break;
            }
        }
        offset++;
        return 0;
    }
};

// binomialCoeffTemplated (common_functionalities.cpp:131-145): Pascal row.
template <class I>
I binom(int n, int k) {
    if (k < 0 || k > n) return 0;
    std::vector<I> Cc(k + 1, (I)0);
    Cc[0] = 1;
    for (int i = 1; i <= n; i++)
        for (int j = std::min(i, k); j > 0; j--) Cc[j] = Cc[j] + Cc[j - 1];
    return Cc[k];
}

template <class T> inline T i128_to(__int128 v) { return T((double)v); }
template <> inline long double i128_to<long double>(__int128 v) { return (long double)v; }
template <> inline Q i128_to<Q>(__int128 v) { return (Q)v; }
template <> inline DD i128_to<DD>(__int128 v) {
    double h = (double)v;
    return DD(h, (double)(v - (__int128)h));
}

// ----------------------------------------------------- BBFG repeated-row --

// BBFGPermanentCalculatorRepeated_Tasks (BBFGPermanentCalculatorRepeated.hpp:
// 74-101 constructor with the anchor row, 116-307 calculate).  col_mult =
// input state (columns), row_mult = output state (rows), as
// BBFGPermanentCalculatorRepeated::calculate(mtx, col_mult, row_mult) (.h:56).
// Returns 0, or 1 with `err` set when the reference throws.
template <class T>
int bbfg_repeated(const double* Ad, int rows_in, int C, int64_t ld, const int64_t* colm64,
                  const int64_t* rowm64, int64_t njobs, int nthreads, Cx<T>* result,
                  std::string* err) {
    if (rows_in == 0) {  // .cpp:78-81
        *result = Cx<T>(T(1.0), T(0.0));
        return 0;
    }
    std::vector<int> row_in(rows_in), col_mult(C);
    for (int i = 0; i < rows_in; ++i) row_in[i] = (int)rowm64[i];  // .cpp:83-92
    for (int j = 0; j < C; ++j) col_mult[j] = (int)colm64[j];
    std::vector<Cx<T>> mtx_in = load<T>(Ad, rows_in, C, ld);

    // constructor, :74-101: anchor = smallest nonzero row multiplicity
    // (first occurrence), comparator of :78
    int minpos = 0;
    for (int i = 1; i < rows_in; ++i) {
        const int a = row_in[i], b = row_in[minpos];
        const bool less = (a == b) ? false : (b == 0 || (a != 0 && a < b));
        if (less) minpos = i;
    }
    std::vector<int> row_mult;
    std::vector<Cx<T>> mtx;
    int R;
    if (rows_in > 0 && row_in[minpos] != 0) {
        R = rows_in + 1;
        row_mult.assign(R, 0);
        row_mult[0] = 1;
        for (int i = 0; i < rows_in; ++i) row_mult[i + 1] = row_in[i];
        row_mult[1 + minpos]--;
        mtx.resize((size_t)R * C);
        for (int c = 0; c < C; ++c) mtx[c] = mtx_in[(size_t)minpos * C + c];
        for (size_t k = 0; k < mtx_in.size(); ++k) mtx[C + k] = mtx_in[k];
    } else {
        R = rows_in;
        row_mult = row_in;
        mtx = mtx_in;
    }

    // calculate(), :116-307
    int sr = 0, sc = 0;
    for (int v : row_mult) sr += v;
    for (int v : col_mult) sc += v;
    if (sr != sc) {
        *err = "BBFGPermanentCalculatorRepeated_Tasks::calculate:  Number of input and output states should be equal";
        return 1;
    }
    if (R == 0 || C == 0 || sr == 0 || sc == 0) {
        *result = Cx<T>(T(1.0), T(0.0));
        return 0;
    }
    if (R == 1) {  // :130-140
        Cx<T> ret(T(1.0), T(0.0));
        for (int idx = 0; idx < C; ++idx)
            for (int j = 0; j < col_mult[idx]; ++j) ret *= mtx[idx];
        *result = ret;
        return 0;
    }
    std::vector<Cx<T>> mtx2(mtx.size());
    for (size_t i = 0; i < mtx.size(); ++i) mtx2[i] = scale(mtx[i], T(2.0));
    std::vector<int> limits(R - 1);
    for (int i = 0; i < R - 1; ++i) limits[i] = row_mult[i + 1] + 1;  // :159-162
    uint64_t Idx_max = (uint64_t)limits[0];
    for (size_t i = 1; i < limits.size(); ++i) Idx_max *= (uint64_t)limits[i];
    int64_t concurrency = std::min<int64_t>(njobs, (int64_t)Idx_max);
    if (concurrency < 1) concurrency = 1;
    std::vector<Cx<T>> part((size_t)concurrency);
    std::vector<int> badflag((size_t)concurrency, 0);
    par_for(concurrency, nthreads, [&](int64_t job_idx) {
        const int64_t work_batch = (int64_t)Idx_max / concurrency;
        const int64_t initial_offset = job_idx * work_batch;
        int64_t offset_max = (job_idx + 1) * work_batch - 1;
        if (job_idx == concurrency - 1) offset_max = (int64_t)Idx_max - 1;
        NaryGray gc(limits, initial_offset);
        gc.offset_max = offset_max;  // set_offset_max
        if (gc.bad) {
            badflag[(size_t)job_idx] = 1;
            return;
        }
        std::vector<int> gcode = gc.gray;
        __int128 bc = 1;
        std::vector<Cx<T>> colsum(mtx.begin(), mtx.begin() + C);
        int minus_all = 0;
        for (size_t idx = 0; idx < gcode.size(); ++idx) {  // :209-226
            const int ms = gcode[idx];
            const int rm = row_mult[idx + 1];
            const T f = T((double)(rm - 2 * ms));
            for (int c = 0; c < C; ++c) {
                const Cx<T>& a = mtx[(idx + 1) * C + c];
                colsum[c] += Cx<T>(a.re * f, a.im * f);
            }
            minus_all += ms;
            bc *= binom<__int128>(rm, ms);
        }
        int parity = (minus_all % 2 == 0) ? 1 : -1;
        Cx<T> prod(T((double)parity), T(0.0));
        for (int c = 0; c < C; ++c)
            for (int j = 0; j < col_mult[c]; ++j) prod *= colsum[c];
        Cx<T> addend;
        {
            const T w = i128_to<T>(bc);
            addend += Cx<T>(prod.re * w, prod.im * w);
        }
        for (int64_t idx = initial_offset + 1; idx < offset_max + 1; ++idx) {  // :248-288
            int changed, vprev, val;
            if (gc.next(changed, vprev, val)) break;
            parity = -parity;
            const Cx<T>* row2 = &mtx2[(size_t)(changed + 1) * C];
            Cx<T> p(T((double)parity), T(0.0));
            for (int c = 0; c < C; ++c) {
                if (vprev < val) colsum[c] -= row2[c];
                else colsum[c] += row2[c];
                for (int j = 0; j < col_mult[c]; ++j) p *= colsum[c];
            }
            const int rm = row_mult[changed + 1];
            bc = val < vprev ? bc * vprev / (rm - val) : bc * (rm - vprev) / val;  // :280
            const T w = i128_to<T>(bc);
            addend += Cx<T>(p.re * w, p.im * w);
        }
        part[(size_t)job_idx] = addend;
    });
    for (int f : badflag)
        if (f) {
            *err = "n_aryGrayCodeCounter::initialize:  Wrong value of initial_offset";
            return 1;
        }
    Cx<T> perm;
    for (const auto& v : part) perm += v;
    const T div = T(std::ldexp(1.0, sr - 1));  // :304
    perm.re = perm.re / div;
    perm.im = perm.im / div;
    *result = perm;
    return 0;
}

// ---------------------------------------------------- Glynn (recursive) --

// GlynnPermanentCalculatorTask<matrix_type, precision>::calculate +
// IterateOverDeltas (GlynnPermanentCalculator.hpp:47-201), serial order.
template <class T>
struct GlynnRec {
    int R, C;
    std::vector<Cx<T>> mtx2;
    ComplexM<T> acc;
    void iterate(std::vector<Cx<T>>& colsum, int sign, int index_min) {
        Cx<T> prod(T(1.0), T(0.0));
        for (int c = 0; c < C; ++c) prod = prod * colsum[c];  // :153-156
        acc += Cx<T>(prod.re * T((double)sign), prod.im * T((double)sign));
        for (int idx = index_min; idx < R; ++idx) {  // :163-181
            std::vector<Cx<T>> cn = colsum;
            for (int j = 0; j < C; ++j) cn[j] = cn[j] - mtx2[(size_t)idx * C + j];
            iterate(cn, -sign, idx + 1);
        }
    }
};

template <class T>
Cx<T> glynn_recursive(const double* Ad, int R, int C, int64_t ld) {
    if (R == 0 || C == 0) return Cx<T>(T(1.0), T(0.0));  // :50-51
    if (R >= C + 2) return Cx<T>(T(0.0), T(0.0));        // :52-53
    GlynnRec<T> g;
    g.R = R;
    g.C = C;
    std::vector<Cx<T>> mtx = load<T>(Ad, R, C, ld);
    g.mtx2.resize(mtx.size());
    for (size_t i = 0; i < mtx.size(); ++i) g.mtx2[i] = scale(mtx[i], T(2.0));  // :80-93
    std::vector<Cx<T>> colsum(C);
    for (int c = 0; c < C; ++c)
        for (int r = 0; r < R; ++r) colsum[c] += mtx[(size_t)r * C + c];  // :103-113
    g.iterate(colsum, 1, 1);
    Cx<T> perm = g.acc.get();
    const T div = T((double)(1ULL << (R - 1)));  // :130
    perm.re = perm.re / div;
    perm.im = perm.im / div;
    return perm;
}

// GlynnPermanentCalculatorRepeated (GlynnPermanentCalculatorRepeated.hpp:
// 49-258); row multiplicities = output state, column multiplicities = input.
template <class T>
struct GlynnRep {
    int R, C;
    std::vector<Cx<T>> mtx2;
    std::vector<int> deltaLimits, colm;
    ComplexM<T> acc;
    void iterate(const std::vector<Cx<T>>& colsum, int sign, int index_min, int64_t mult) {
        Cx<T> prod(T(1.0), T(0.0));
        for (int c = 0; c < C; ++c)
            for (int j = 0; j < colm[c]; ++j) prod = prod * colsum[c];  // :214-219
        const T f = T((double)(mult * sign));
        acc += Cx<T>(prod.re * f, prod.im * f);
        for (int idx = index_min; idx < R; ++idx) {  // :229-255
            int localSign = sign;
            std::vector<Cx<T>> cn = colsum;
            for (int m = 1; m <= deltaLimits[idx]; ++m) {
                for (int j = 0; j < C; ++j) cn[j] = cn[j] - mtx2[(size_t)idx * C + j];
                localSign *= -1;
                iterate(cn, localSign, idx + 1, mult * binom<int64_t>(deltaLimits[idx], m));
            }
        }
    }
};

template <class T>
int glynn_repeated(const double* Ad, int R, int C, int64_t ld, const int64_t* input_state,
                   const int64_t* output_state, Cx<T>* result, std::string* err) {
    int64_t si = 0, so = 0;
    for (int j = 0; j < C; ++j) si += input_state[j];
    for (int i = 0; i < R; ++i) so += output_state[i];
    if (si != so) {  // :55-60
        *err = "GlynnPermanentCalculatorRepeated::calculate:  Number of input and output states should be equal";
        return 1;
    }
    if (R == 0 || si == 0 || so == 0) {  // :62-63
        *result = Cx<T>(T(1.0), T(0.0));
        return 0;
    }
    GlynnRep<T> g;
    g.R = R;
    g.C = C;
    std::vector<Cx<T>> mtx = load<T>(Ad, R, C, ld);
    g.mtx2.resize(mtx.size());
    for (size_t i = 0; i < mtx.size(); ++i) g.mtx2[i] = scale(mtx[i], T(2.0));
    std::vector<int> rowm(R);
    g.colm.resize(C);
    for (int i = 0; i < R; ++i) rowm[i] = (int)output_state[i];
    for (int j = 0; j < C; ++j) g.colm[j] = (int)input_state[j];
    int minimalIndex = R;  // :112-130
    g.deltaLimits.assign(R, 0);
    for (int i = 0; i < R; ++i) {
        if (rowm[i] > 0) {
            if (minimalIndex > i) {
                g.deltaLimits[i] = rowm[i] - 1;
                minimalIndex = i;
            } else {
                g.deltaLimits[i] = rowm[i];
            }
        }
    }
    if (minimalIndex == R) {  // :142-144
        *result = Cx<T>(T(1.0), T(0.0));
        return 0;
    }
    std::vector<Cx<T>> colsum(C);
    for (int c = 0; c < C; ++c)
        for (int r = 0; r < R; ++r) {
            const T f = T((double)rowm[r]);
            const Cx<T>& a = mtx[(size_t)r * C + c];
            colsum[c] += Cx<T>(f * a.re, f * a.im);  // :153-163
        }
    g.iterate(colsum, 1, minimalIndex, 1);
    Cx<T> perm = g.acc.get();
    const T div = T((double)(1ULL << (so - 1)));  // :185-190
    perm.re = perm.re / div;
    perm.im = perm.im / div;
    *result = perm;
    return 0;
}

// GlynnPermanentCalculatorSimpleTask::calculate (GlynnPermanentCalculatorSimple.hpp:61-168),
// serial: one range [0, 2^(n-1)).
template <class T>
Cx<T> glynn_simple(const double* Ad, int R, int C, int64_t ld) {
    if (R == 0 || C == 0) return Cx<T>(T(1.0), T(0.0));
    if (R != C) return Cx<T>(T(0.0), T(0.0));
    const int n = R;
    std::vector<Cx<T>> mtx = load<T>(Ad, R, C, ld);
    const uint64_t powerOf2 = 1ULL << (n - 1);
    std::vector<int> deltas(n, 1);
    int sign = 1;  // range.begin() == 0: all deltas +1
    ComplexM<T> acc;
    for (uint64_t d = 0; d < powerOf2; ++d) {
        Cx<T> product(T(1.0), T(0.0));
        for (int r = 0; r < n; ++r) {
            Cx<T> rowSum = mtx[(size_t)r * C];
            for (int c = 1; c < n; ++c) {
                const T f = T((double)deltas[c]);
                const Cx<T>& a = mtx[(size_t)r * C + c];
                rowSum += Cx<T>(f * a.re, f * a.im);
            }
            product *= rowSum;
        }
        acc += Cx<T>(product.re * T((double)sign), product.im * T((double)sign));
        uint64_t nextD = d + 1;
        int indexChanged = 1;
        while (nextD > 0) {
            if (nextD % 2 == 1) nextD = 0;
            else {
                nextD >>= 1;
                ++indexChanged;
            }
        }
        if (indexChanged % 2 != 0) sign *= -1;
        for (int k = 0; k < indexChanged; ++k) deltas[n - k - 1] *= -1;
    }
    Cx<T> perm = acc.get();
    const T div = T((double)powerOf2);
    perm.re = perm.re / div;
    perm.im = perm.im / div;
    return perm;
}

// ====================================================== independent checks ==

// Sum over all permutations (n! terms); n <= 10.
template <class T>
Cx<T> perm_bruteforce(const double* Ad, int n, int64_t ld) {
    if (n == 0) return Cx<T>(T(1.0), T(0.0));
    std::vector<Cx<T>> m = load<T>(Ad, n, n, ld);
    std::vector<int> p(n);
    for (int i = 0; i < n; ++i) p[i] = i;
    Cx<T> s;
    // Heap's algorithm
    std::vector<int> c(n, 0);
    auto term = [&]() {
        Cx<T> t(T(1.0), T(0.0));
        for (int i = 0; i < n; ++i) t *= m[(size_t)i * n + p[i]];
        s += t;
    };
    term();
    int i = 0;
    while (i < n) {
        if (c[i] < i) {
            if (i % 2 == 0) std::swap(p[0], p[i]);
            else std::swap(p[c[i]], p[i]);
            term();
            c[i]++;
            i = 0;
        } else {
            c[i] = 0;
            i++;
        }
    }
    return s;
}

// Ryser: perm = (-1)^n sum_{S subset of cols} (-1)^|S| prod_i sum_{j in S} a_ij,
// subsets visited in binary Gray order; parallel over the top bits.
template <class T>
Cx<T> perm_ryser(const double* Ad, int n, int64_t ld, int nthreads) {
    if (n == 0) return Cx<T>(T(1.0), T(0.0));
    std::vector<Cx<T>> m = load<T>(Ad, n, n, ld);
    const int top = std::min(n, 6);
    const int low = n - top;
    const int64_t njobs = (int64_t)1 << top;
    std::vector<Cx<T>> part((size_t)njobs);
    par_for(njobs, nthreads, [&](int64_t job) {
        // columns [0, low) walked by Gray code, columns [low, n) fixed by `job`
        std::vector<Cx<T>> rs(n);
        int cnt = 0;
        for (int b = 0; b < top; ++b)
            if ((job >> b) & 1) {
                ++cnt;
                for (int i = 0; i < n; ++i) rs[i] += m[(size_t)i * n + low + b];
            }
        Cx<T> acc;
        const uint64_t W = 1ULL << low;
        uint64_t g = 0;
        for (uint64_t t = 0; t < W; ++t) {
            if (t > 0) {
                const int b = __builtin_ctzll(t);
                const uint64_t ng = t ^ (t >> 1);
                const bool add = (ng >> b) & 1;
                for (int i = 0; i < n; ++i) {
                    if (add) rs[i] += m[(size_t)i * n + b];
                    else rs[i] -= m[(size_t)i * n + b];
                }
                cnt += add ? 1 : -1;
                g = ng;
            }
            if (cnt == 0) continue;  // empty set contributes prod of zeros = 0 (n >= 1)
            Cx<T> pr = rs[0];
            for (int i = 1; i < n; ++i) pr *= rs[i];
            if (((n - cnt) & 1) == 0) acc += pr; else acc -= pr;
        }
        (void)g;
        part[(size_t)job] = acc;
    });
    Cx<T> s;
    for (auto& v : part) s += v;
    return s;
}

// Exact Ryser on an integer matrix; every product is computed in signed
// 128-bit without overflow when prod_i sum_j |a_ij| < 2^126 (checked).
int perm_ryser_int128(const int64_t* A, int n, int64_t ld, int nthreads, __int128* out) {
    if (n == 0) {
        *out = 1;
        return 0;
    }
    long double lg = 0.0L;
    for (int i = 0; i < n; ++i) {
        long double rs = 0.0L;
        for (int j = 0; j < n; ++j) rs += std::fabs((long double)A[i * ld + j]);
        if (rs == 0.0L) {
            *out = 0;
            return 0;
        }
        lg += std::log2(rs);
    }
    if (lg >= 126.0L) return 1;
    const int top = std::min(n, 6);
    const int low = n - top;
    const int64_t njobs = (int64_t)1 << top;
    // sums are kept modulo 2^128 (unsigned, wrap-around is defined); the true
    // permanent is below 2^126 in magnitude, so the final value is exact.
    std::vector<unsigned __int128> part((size_t)njobs, 0);
    par_for(njobs, nthreads, [&](int64_t job) {
        std::vector<int64_t> rs(n, 0);
        int cnt = 0;
        for (int b = 0; b < top; ++b)
            if ((job >> b) & 1) {
                ++cnt;
                for (int i = 0; i < n; ++i) rs[i] += A[i * ld + low + b];
            }
        unsigned __int128 acc = 0;
        const uint64_t W = 1ULL << low;
        for (uint64_t t = 0; t < W; ++t) {
            if (t > 0) {
                const int b = __builtin_ctzll(t);
                const uint64_t ng = t ^ (t >> 1);
                const bool add = (ng >> b) & 1;
                for (int i = 0; i < n; ++i) rs[i] += add ? A[i * ld + b] : -A[i * ld + b];
                cnt += add ? 1 : -1;
            }
            if (cnt == 0) continue;
            __int128 pr = rs[0];
            for (int i = 1; i < n; ++i) pr *= rs[i];
            if (((n - cnt) & 1) == 0) acc += (unsigned __int128)pr; else acc -= (unsigned __int128)pr;
        }
        part[(size_t)job] = acc;
    });
    unsigned __int128 s = 0;
    for (auto v : part) s += v;
    *out = (__int128)s;
    return 0;
}

// ---------------------------------------------------------------- dispatch --

template <class F>
int by_prec(int prec, F&& f) {
    switch (prec) {
        case 0: return f(double());
        case 1: return f((long double)0);
        case 2: return f(DD());
        case 3: return f(Q());
        default: return -1;
    }
}

thread_local std::string g_err;

}  // namespace

extern "C" {

// prec: 0 double, 1 long double (x87), 2 double-double, 3 binary128.
// out: 6 doubles [re, im, hi_re, lo_re, hi_im, lo_im].

const char* orc_last_error(void) { return g_err.c_str(); }

int orc_bbfg(const double* A, int rows, int cols, int64_t ld, int prec, int64_t njobs, int nthreads,
             double* out) {
    return by_prec(prec, [&](auto z) {
        using T = decltype(z);
        store(bbfg_plain<T>(A, rows, cols, ld, njobs, nthreads), out);
        return 0;
    });
}

// Partial sweep (jobs [job_lo, job_hi) of `njobs`), used to time a bounded
// sample of the reference path.  Returns the (unnormalised) partial sum / 2^(R-1).
int orc_bbfg_partial(const double* A, int rows, int cols, int64_t ld, int prec, int64_t njobs,
                     int64_t job_lo, int64_t job_hi, int nthreads, double* out) {
    return by_prec(prec, [&](auto z) {
        using T = decltype(z);
        store(bbfg_plain<T>(A, rows, cols, ld, njobs, nthreads, job_lo, job_hi), out);
        return 0;
    });
}

int orc_bbfg_repeated(const double* A, int rows, int cols, int64_t ld, const int64_t* row_mult,
                      const int64_t* col_mult, int prec, int64_t njobs, int nthreads, double* out) {
    return by_prec(prec, [&](auto z) {
        using T = decltype(z);
        Cx<T> r;
        std::string e;
        if (bbfg_repeated<T>(A, rows, cols, ld, col_mult, row_mult, njobs, nthreads, &r, &e)) {
            g_err = e;
            return 1;
        }
        store(r, out);
        return 0;
    });
}

int orc_glynn_recursive(const double* A, int rows, int cols, int64_t ld, int prec, double* out) {
    return by_prec(prec, [&](auto z) {
        using T = decltype(z);
        store(glynn_recursive<T>(A, rows, cols, ld), out);
        return 0;
    });
}

int orc_glynn_repeated(const double* A, int rows, int cols, int64_t ld, const int64_t* input_state,
                       const int64_t* output_state, int prec, double* out) {
    return by_prec(prec, [&](auto z) {
        using T = decltype(z);
        Cx<T> r;
        std::string e;
        if (glynn_repeated<T>(A, rows, cols, ld, input_state, output_state, &r, &e)) {
            g_err = e;
            return 1;
        }
        store(r, out);
        return 0;
    });
}

int orc_glynn_simple(const double* A, int rows, int cols, int64_t ld, int prec, double* out) {
    return by_prec(prec, [&](auto z) {
        using T = decltype(z);
        store(glynn_simple<T>(A, rows, cols, ld), out);
        return 0;
    });
}

int orc_perm_bruteforce(const double* A, int n, int64_t ld, int prec, double* out) {
    if (n > 11) return -2;
    return by_prec(prec, [&](auto z) {
        using T = decltype(z);
        store(perm_bruteforce<T>(A, n, ld), out);
        return 0;
    });
}

int orc_perm_ryser(const double* A, int n, int64_t ld, int prec, int nthreads, double* out) {
    if (n > 40) return -2;
    return by_prec(prec, [&](auto z) {
        using T = decltype(z);
        store(perm_ryser<T>(A, n, ld, nthreads), out);
        return 0;
    });
}

// words[0] = low 64 bits, words[1] = high 64 bits of the signed permanent.
int orc_perm_int128(const int64_t* A, int n, int64_t ld, int nthreads, int64_t* words) {
    __int128 v = 0;
    if (perm_ryser_int128(A, n, ld, nthreads, &v)) {
        g_err = "integer permanent may exceed the signed 128-bit range";
        return 1;
    }
    words[0] = (int64_t)(unsigned long long)v;
    words[1] = (int64_t)(v >> 64);
    return 0;
}

int orc_hardware_threads(void) { return (int)std::thread::hardware_concurrency(); }

}  // extern "C"
