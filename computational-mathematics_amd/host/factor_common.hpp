// Shared host-side structures for the factor producers (AFIIM, ACIFIM, PDFIIM).
//
// The producers run once per system on the host; their output is the factor
// interface consumed by the GPU solve (CSR, 0-based, int32 indices, columns
// ascending; G unit upper with the diagonal first in each row, H unit lower
// with the diagonal last, IDv = D^{-1}).  Layout contract:
// AFIIM/src/prec.cpp:682-740, PDFIIM/src/parPrec.cpp:536-587,
// ACIFIM/src/prec.cpp:900-1011.
#pragma once
#include <cstdint>
#include <vector>

namespace cmh {

// One stored entry of a growing row or column list.
struct Ent {
    int i;
    double v;
};

// Sparse accumulator over [0, n) with a singly linked list of the touched
// positions.  New positions are pushed at the head, and every traversal
// starts at the head, so positions come back most-recent-first.  The factor
// recursions depend on that order for their floating-point sums, so it is
// part of the contract (AFIIM/src/prec.cpp:365-370,388-393).
struct Spa {
    std::vector<double> v;
    std::vector<int> link;  // -1: not present; otherwise next position (-2 ends)
    int head = -2;
    int cnt = 0;

    void init(int n)
    {
        v.assign(n, 0.0);
        link.assign(n, -1);
        head = -2;
        cnt = 0;
    }
    // First touch stores x, later touches accumulate.
    inline void add(int k, double x)
    {
        if (link[k] == -1) {
            v[k] = x;
            link[k] = head;
            head = k;
            cnt++;
        } else {
            v[k] += x;
        }
    }
    template <class F>
    inline void walk(F f) const
    {
        int k = head;
        for (int j = 0; j < cnt; j++) {
            int nx = link[k];
            f(k);
            k = nx;
        }
    }
    inline void clear()
    {
        int k = head;
        for (int j = 0; j < cnt; j++) {
            int nx = link[k];
            link[k] = -1;
            v[k] = 0.0;
            k = nx;
        }
        head = -2;
        cnt = 0;
    }
};

// Strictly-lower part of A by rows (A's column order) and strictly-upper part
// of A by columns (row order), plus the diagonal: the split used by the
// bordering recursions (AFIIM/src/prec.cpp:239-306).
struct Split {
    std::vector<int> lptr, lidx;
    std::vector<double> lval;
    std::vector<int> uptr, uidx;  // column-wise: uidx holds row indices
    std::vector<double> uval;
    std::vector<double> d;
};

void split_lower_upper(int n, const double *Av, const int *Aj, const int *Ai,
                       double zero_diag_value, Split &s);

inline double dmax(double a, double b) { return (a > b) ? a : b; }
inline double dmin(double a, double b) { return (a < b) ? a : b; }

// Output arrays are allocated with new[] so that a caller written against the
// reference interface can release them with delete[] exactly as before
// (AFIIM/example/example.cpp:127-137).
template <class T>
T *new_copy(const std::vector<T> &v)
{
    T *p = new T[v.size() > 0 ? v.size() : 1];
    for (size_t k = 0; k < v.size(); k++) p[k] = v[k];
    return p;
}

}  // namespace cmh
