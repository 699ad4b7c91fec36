// PDFIIM factor construction on the GPU (PDFIIM/src/parPrec.cpp:95-588).
//
// Row i of the construction is independent of every other row: a sparsity
// pattern S of indices < i grown by level sets of A's graph, the small dense
// systems g_S = -A(S,S)^{-1} A(S,i) and h_S = -A(S,S)^{-T} A(i,S)^T solved
// with partial pivoting, the solutions filtered against a norm, S rebuilt
// from the survivors and grown again, lfill times; then column i of G, row i
// of H and D(i) = 1 / ([h;1]^T A [g;1]).  One wavefront (one workgroup of
// 64 lanes) builds one row; the grid strides over the rows.
//
// The pattern is kept as the reference keeps it (PDFIIM/src/
// sparseAccumulator.hpp:268-482): an insertion-ordered set whose traversal
// runs most-recent-first, and whose insertion order fixes the row/column
// order of the dense system.  Candidates of one source row are tested by
// all lanes at once (each against the set and against the earlier lanes'
// candidates) and appended in order, which is the reference's sequential
// push.  The dense system lives in LDS, column-major with a padded stride,
// one lane per row: the LU sweeps the columns, the pivot search is a wave
// reduction with the lowest index winning ties, and the triangular solves
// are column-oriented, so every scalar goes through the same operations in
// the same order as the host producer (host/pdfiim.cpp), with no fused
// multiply-adds: the factors are the host producer's bit for bit.
//
// A pattern longer than CAP at a solve (or a set longer than LIST at any
// time) flags the row; the host re-runs flagged rows with the next CAP.
#pragma once
#include "cm_kernels.hpp"

namespace cm {

constexpr int PD_WAVE = 64;

struct PdParams {
    int n, prefill, lfill, postfill, warpsize;
    double ftol, diag_floor;
    char filttype, patt;
};

struct PdCsr {
    const int *ptr, *idx;
    const double *val;
};

// Per-row results: counts, the sorted survivors of g (column i of G, the unit
// diagonal last) and of h (row i of H, the unit diagonal last) and D(i);
// counts, flags and D by row, the entries in slots of `out` per listed row
// (slot q of the row list, or row i - 1 without a list).
struct PdOut {
    int *gcnt, *hcnt, *flag;
    int *gidx, *hidx;
    double *gval, *hval, *dinv;
    int out;
};

template <int CAP>
struct PdShared {
    static constexpr int LD = CAP + 1;      // padded column stride of B
    static constexpr int LIST = 4 * CAP;    // pattern set capacity while it grows
    double B[CAP * LD];
    double gsv[CAP + 1], hsv[CAP + 1];
    int list[LIST];
    int inds[CAP];
    int gsi[CAP + 1], hsi[CAP + 1];
    int cand[PD_WAVE];
};

__device__ __forceinline__ void pd_sync() { __syncthreads(); }

__device__ __forceinline__ double pd_shfl(double v, int src) { return __shfl(v, src, PD_WAVE); }
__device__ __forceinline__ int pd_shfl(int v, int src) { return __shfl(v, src, PD_WAVE); }

__device__ __forceinline__ unsigned long long pd_ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int pd_below(unsigned long long mask)
{
    const int lane = threadIdx.x;
    return __popcll(lane ? (mask & ((1ull << lane) - 1ull)) : 0ull);
}

// Entries of row r of m with index < lim (the rows are sorted: a prefix).
__device__ __forceinline__ int pd_prefix(const PdCsr &m, int r, int lim)
{
    const int s = m.ptr[r], e = m.ptr[r + 1];
    int cnt = 0;
    for (int b = s; b < e; b += PD_WAVE) {
        const int k = b + (int)threadIdx.x;
        const bool below = k < e && m.idx[k] < lim;
        const int c = __popcll(pd_ballot(below));
        cnt += c;
        if (c < PD_WAVE) break;
    }
    return cnt;
}

template <int CAP>
struct PdRow {
    PdShared<CAP> &sh;
    const PdParams &P;
    const PdCsr &A, &At;
    int i;
    int cnt = 0;       // pattern set size (uniform)
    bool ovf = false;  // pattern beyond CAP / LIST: the row goes to the next tier

    __device__ PdRow(PdShared<CAP> &s, const PdParams &p, const PdCsr &a, const PdCsr &at, int row)
        : sh(s), P(p), A(a), At(at), i(row)
    {
    }

    // push(c) for the entries < i of At row r, then of A row r, in order
    // (PDFIIM/src/parPrec.cpp:150-170): each lane tests one candidate
    // against the set and the earlier lanes' candidates; the new ones are
    // appended in lane order
    __device__ void push_rows(int r)
    {
        const bool ua = P.patt == 'U' || P.patt == 'B', ul = P.patt == 'L' || P.patt == 'B';
        const int na = ua ? pd_prefix(At, r, i) : 0;
        const int nb = ul ? pd_prefix(A, r, i) : 0;
        const int lane = threadIdx.x;
        for (int base = 0; base < na + nb; base += PD_WAVE) {
            const int q = base + lane;
            const bool have = q < na + nb;
            const int c = have ? (q < na ? At.idx[At.ptr[r] + q] : A.idx[A.ptr[r] + q - na]) : -1;
            sh.cand[lane] = c;
            pd_sync();
            bool isnew = have;
            for (int t = 0; t < lane && isnew; t++)
                if (sh.cand[t] == c) isnew = false;
            const int lim = cnt < PdShared<CAP>::LIST ? cnt : PdShared<CAP>::LIST;
            for (int t = 0; t < lim && isnew; t++)
                if (sh.list[t] == c) isnew = false;
            const unsigned long long m = pd_ballot(isnew);
            const int pos = cnt + pd_below(m);
            if (isnew && pos < PdShared<CAP>::LIST) sh.list[pos] = c;
            cnt += __popcll(m);
            if (cnt > PdShared<CAP>::LIST) ovf = true;
            pd_sync();
            if (ovf) return;
        }
    }

    // the rows of the set's current members, most recent first (rewind +
    // next over the snapshot; later pushes are not visited)
    __device__ void expand_snapshot()
    {
        const int len = cnt;
        for (int t = len - 1; t >= 0 && !ovf; t--) push_rows(sh.list[t]);
    }

    // rank of index c in the set's first len members, -1 if absent
    __device__ int rank_of(int c, int len) const
    {
        for (int t = 0; t < len; t++)
            if (sh.inds[t] == c) return t;
        return -1;
    }

    __device__ double &Bel(int r, int c) { return sh.B[c * PdShared<CAP>::LD + r]; }

    // one solve: LU of A(S,S) (S = inds[0..len)), g = B^{-1} rhs, h = B^{-T} rhs2;
    // lane r returns g_r in v and h_r in w
    __device__ void solve(int len, double &v, double &w)
    {
#pragma clang fp contract(off)
        const int lane = threadIdx.x;
        const bool mine = lane < len;
        // B(r, :) = A(inds[r], inds[.]) restricted to columns < i, floor on the diagonal
        // (PDFIIM/src/parPrec.cpp:252-270)
        if (mine) {
            for (int c = 0; c < len; c++) Bel(lane, c) = 0.0;
            Bel(lane, lane) = P.diag_floor;
            const int r = sh.inds[lane];
            for (int kk = A.ptr[r]; kk < A.ptr[r + 1] && A.idx[kk] < i; kk++) {
                const int k2 = rank_of(A.idx[kk], len);
                if (k2 >= 0) Bel(lane, k2) = A.val[kk];
            }
            if (Bel(lane, lane) == 0.0) Bel(lane, lane) = P.diag_floor;
        }
        // rhs = -A(S, i) from column i of A (row i of A^T), rhs2 = -A(i, S)
        // (PDFIIM/src/parPrec.cpp:272-284)
        if (lane < CAP) {
            sh.gsv[lane] = 0.0;
            sh.hsv[lane] = 0.0;
        }
        pd_sync();
        for (int q = At.ptr[i] + lane; q < At.ptr[i + 1]; q += PD_WAVE) {
            const int c = At.idx[q];
            if (c < i) {
                const int k2 = rank_of(c, len);
                if (k2 >= 0) sh.gsv[k2] = -At.val[q];
            }
        }
        for (int q = A.ptr[i] + lane; q < A.ptr[i + 1]; q += PD_WAVE) {
            const int c = A.idx[q];
            if (c < i) {
                const int k2 = rank_of(c, len);
                if (k2 >= 0) sh.hsv[k2] = -A.val[q];
            }
        }
        pd_sync();
        v = mine ? sh.gsv[lane] : 0.0;
        w = mine ? sh.hsv[lane] : 0.0;
        int piv_mine = lane;  // lane j keeps piv[j]
        // LU with partial pivoting, row-major semantics of the host's lu_factor
        for (int j = 0; j < len; j++) {
            const double ajj = Bel(j, j);
            double key = -1.0;
            if (lane >= j && mine) {
                const double a = fabs(Bel(lane, j));
                key = (a == a) ? a : -1.0;  // NaN is never chosen over an earlier value
            }
            int p = lane;
            // wave argmax, lowest index on ties
            for (int off = 32; off >= 1; off >>= 1) {
                const double ko = __shfl_xor(key, off, PD_WAVE);
                const int po = __shfl_xor(p, off, PD_WAVE);
                if (ko > key || (ko == key && po < p)) {
                    key = ko;
                    p = po;
                }
            }
            if (!(fabs(ajj) == fabs(ajj))) p = j;  // NaN pivot candidate: the host keeps j
            else if (!(key > fabs(ajj))) p = j;     // nothing strictly larger than |B(j,j)|
            if (lane == j) piv_mine = p;
            pd_sync();
            if (p != j && mine) {  // swap rows j and p: lane c swaps column c's two entries
                const double a = Bel(j, lane), b = Bel(p, lane);
                Bel(j, lane) = b;
                Bel(p, lane) = a;
            }
            pd_sync();
            const double d = Bel(j, j);
            if (d != 0.0 && mine && lane > j) {
                const double l = Bel(lane, j) / d;
                Bel(lane, j) = l;
                if (l != 0.0)
                    for (int c = j + 1; c < len; c++) Bel(lane, c) = Bel(lane, c) - l * Bel(j, c);
            }
            pd_sync();
        }
        // B x = rhs: the row swaps in order, then L (unit) and U column by column
        for (int j = 0; j < len; j++) {
            const int pj = pd_shfl(piv_mine, j);
            if (pj != j) {
                const double vj = pd_shfl(v, j), vp = pd_shfl(v, pj);
                if (lane == j) v = vp;
                if (lane == pj) v = vj;
            }
        }
        for (int c = 0; c < len; c++) {
            const double bc = pd_shfl(v, c);
            if (mine && lane > c) v = v - Bel(lane, c) * bc;
        }
        for (int c = len - 1; c >= 0; c--) {
            if (lane == c) v = v / Bel(c, c);
            const double bc = pd_shfl(v, c);
            if (lane < c) v = v - Bel(lane, c) * bc;
        }
        // B^T x = rhs2: U^T, then L^T (unit), then the row swaps in reverse
        for (int c = 0; c < len; c++) {
            if (lane == c) w = w / Bel(c, c);
            const double bc = pd_shfl(w, c);
            if (mine && lane > c) w = w - Bel(c, lane) * bc;
        }
        for (int c = len - 1; c >= 0; c--) {
            const double bc = pd_shfl(w, c);
            if (lane < c) w = w - Bel(c, lane) * bc;
        }
        for (int j = len - 1; j >= 0; j--) {
            const int pj = pd_shfl(piv_mine, j);
            if (pj != j) {
                const double wj = pd_shfl(w, j), wp = pd_shfl(w, pj);
                if (lane == j) w = wp;
                if (lane == pj) w = wj;
            }
        }
        pd_sync();
    }

    // the filter norm of the reference (PDFIIM/src/parPrec.cpp:293-323),
    // summed in lane order
    __device__ double norm(double v, int len) const
    {
#pragma clang fp contract(off)
        double nrm = 1.0;
        if (P.filttype == '1') {
            for (int j = 0; j < len; j++) nrm += fabs(pd_shfl(v, j));
            nrm /= (len + 1);
        } else if (P.filttype == '2') {
            for (int j = 0; j < len; j++) {
                const double x = pd_shfl(v, j);
                nrm += x * x;
            }
            nrm = sqrt(nrm);
        } else {
            for (int j = 0; j < len; j++) {
                const double x = fabs(pd_shfl(v, j));
                nrm = (nrm < x) ? x : nrm;
            }
        }
        return nrm;
    }

    // survivors sorted by index into (si, sv), the unit diagonal i appended
    __device__ int store_sorted(bool surv, double val, int len, int *si, double *sv)
    {
        const int lane = threadIdx.x;
        const int me = lane < len ? sh.inds[lane] : 0;
        const unsigned long long m = pd_ballot(surv);
        int pos = 0;
        for (int t = 0; t < len; t++) {
            const int o = pd_shfl(me, t);
            if (((m >> t) & 1ull) && o < me) pos++;
        }
        if (surv) {
            si[pos] = me;
            sv[pos] = val;
        }
        const int c = __popcll(m);
        if (lane == 0) {
            si[c] = i;
            sv[c] = 1.0;
        }
        pd_sync();
        return c + 1;
    }

    __device__ void run(const PdOut &O, size_t slot)
    {
#pragma clang fp contract(off)
        const int lane = threadIdx.x;
        cnt = 0;
        push_rows(i);  // A^T row i, then A row i (PDFIIM/src/parPrec.cpp:150-158)
        for (int k = 2; k <= P.prefill && !ovf; k++) expand_snapshot();
        double v = 0.0, w = 0.0, ng = 0.0, nh = 0.0;
        int tlen = 0;
        for (int k = 1; k <= P.lfill && !ovf; k++) {
            ng = 0.0;
            int len = cnt;
            tlen = len;
            if (len == 0) break;
            if (P.warpsize != -1 && len >= P.warpsize) cnt = P.warpsize - 1;  // delete_last down to warpsize - 1
            len = cnt;
            tlen = len;
            if (len > CAP) {
                ovf = true;
                break;
            }
            if (lane < len) sh.inds[lane] = sh.list[lane];
            pd_sync();
            solve(len, v, w);
            cnt = 0;
            ng = norm(v, len);
            nh = norm(w, len);
            const bool gs = lane < len && fabs(v) >= P.ftol * ng;
            const bool hs = lane < len && fabs(w) >= P.ftol * nh;
            const unsigned long long mg = pd_ballot(gs), mh = pd_ballot(hs);
            const int c = __popcll(mg), c2 = __popcll(mh);
            // the set again: g's survivors, then h's that are not among them
            const unsigned long long mu = pd_ballot(hs && !gs);
            if (gs) sh.list[pd_below(mg)] = sh.inds[lane];
            if (hs && !gs) sh.list[c + pd_below(mu)] = sh.inds[lane];
            cnt = c + __popcll(mu);
            pd_sync();
            if ((c == 0 && c2 == 0) || c2 == P.warpsize) break;
            if (k == P.lfill) break;  // the next level set only feeds a further pass
            const int len_new = cnt;
            // level set of the first len_new entries of this pass's order
            // (PDFIIM/src/parPrec.cpp:350-371 reads inds, not the survivors)
            for (int j = 0; j < len_new && !ovf; j++) push_rows(sh.inds[j]);
            for (int m2 = 2; m2 <= P.postfill && !ovf; m2++) expand_snapshot();
            if (cnt == len_new) break;
        }
        if (ovf) {
            if (lane == 0) O.flag[i] = 1;
            return;
        }
        const bool gs = lane < tlen && fabs(v) >= P.ftol * ng;
        const bool hs = lane < tlen && fabs(w) >= P.ftol * nh;
        const int gn = store_sorted(gs, v, tlen, sh.gsi, sh.gsv);
        const int hn = store_sorted(hs, w, tlen, sh.hsi, sh.hsv);
        const size_t o = slot * O.out;
        for (int t = lane; t < gn; t += PD_WAVE) {
            O.gidx[o + t] = sh.gsi[t];
            O.gval[o + t] = sh.gsv[t];
        }
        for (int t = lane; t < hn; t += PD_WAVE) {
            O.hidx[o + t] = sh.hsi[t];
            O.hval[o + t] = sh.hsv[t];
        }
        // D(i) = 1 / (h^T A g): for every h entry (sorted), A's row against g
        // in increasing column order (PDFIIM/src/parPrec.cpp:507-532); lanes
        // find the matches of 64 row entries at a time, the sum runs in order
        double s = 0.0;
        for (int q = 0; q < hn; q++) {
            const int hj = sh.hsi[q];
            const double hv = sh.hsv[q];
            for (int b = A.ptr[hj]; b < A.ptr[hj + 1]; b += PD_WAVE) {
                const int kk = b + lane;
                double prod = 0.0;
                bool hit = false;
                if (kk < A.ptr[hj + 1]) {
                    const int col = A.idx[kk];
                    int lo = 0, hi = gn - 1;  // g's indices are sorted
                    while (lo <= hi) {
                        const int mid = (lo + hi) >> 1;
                        const int gm = sh.gsi[mid];
                        if (gm == col) {
                            hit = true;
                            prod = hv * A.val[kk] * sh.gsv[mid];
                            break;
                        }
                        if (gm < col) lo = mid + 1;
                        else hi = mid - 1;
                    }
                }
                const unsigned long long mh2 = pd_ballot(hit);
                for (int t = 0; t < PD_WAVE; t++) {
                    if (!((mh2 >> t) & 1ull)) continue;
                    s += pd_shfl(prod, t);
                }
            }
        }
        if (lane == 0) {
            O.gcnt[i] = gn;
            O.hcnt[i] = hn;
            O.dinv[i] = 1. / s;
            O.flag[i] = 0;
        }
    }
};

// One workgroup (one wavefront) per row; rows[] lists the rows to build
// (all rows 1..n-1 first, then the ones a smaller CAP flagged).
template <int CAP>
__global__ __launch_bounds__(PD_WAVE) void k_pdfiim_rows(int nrows, const int *__restrict__ rows, PdParams P,
                                                         PdCsr A, PdCsr At, PdOut O)
{
    __shared__ PdShared<CAP> sh;
    for (int q = blockIdx.x; q < nrows; q += gridDim.x) {
        const int i = rows ? rows[q] : q + 1;
        PdRow<CAP> row(sh, P, A, At, i);
        row.run(O, (size_t)q);
        __syncthreads();
    }
}

}  // namespace cm

namespace cm {

// Packs the listed rows' slot entries to their offsets (ofs[i]: where row i's
// cnt[i] entries go in the packed arrays).
__global__ __launch_bounds__(BS) void k_pd_pack(int nrows, const int *__restrict__ rows, int out,
                                                const int *__restrict__ cnt, const long long *__restrict__ ofs,
                                                const int *__restrict__ sidx, const double *__restrict__ sval,
                                                int *__restrict__ pidx, double *__restrict__ pval)
{
    for (int q = blockIdx.x * BS + threadIdx.x; q < nrows; q += gridDim.x * BS) {
        const int i = rows ? rows[q] : q + 1;
        // a row rebuilt by a later tier has a longer count than this tier's
        // slot holds: copy no more than the slot (the later tier's pack
        // overwrites the row)
        const int c = cnt[i] < out ? cnt[i] : out;
        const long long o = ofs[i];
        for (int t = 0; t < c; t++) {
            pidx[o + t] = sidx[(size_t)q * out + t];
            pval[o + t] = sval[(size_t)q * out + t];
        }
    }
}

}  // namespace cm

namespace cm {

// The rows whose patterns outgrow 64 indices: the same construction with the
// dense system, the vectors and the set in a per-workgroup slab of global
// memory (up to PD_BIG indices per solve), each lane taking rows lane,
// lane + 64, ...  Every scalar still goes through the host producer's
// operations in its order.
constexpr int PD_BIG = 256;

struct PdBigSlab {
    static constexpr int LD = PD_BIG + 1;
    static constexpr int LIST = 4 * PD_BIG;
    double B[PD_BIG * LD];
    double V[PD_BIG], W[PD_BIG];
    double gsv[PD_BIG + 1], hsv[PD_BIG + 1];
    int list[LIST];
    int inds[PD_BIG], piv[PD_BIG];
    int gsi[PD_BIG + 1], hsi[PD_BIG + 1];
    int GS[PD_BIG], HS[PD_BIG];
    int cand[PD_WAVE];
};

struct PdRowBig {
    PdBigSlab &sl;
    const PdParams &P;
    const PdCsr &A, &At;
    int i;
    int cnt = 0;
    bool ovf = false;

    __device__ PdRowBig(PdBigSlab &s, const PdParams &p, const PdCsr &a, const PdCsr &at, int row)
        : sl(s), P(p), A(a), At(at), i(row)
    {
    }

    __device__ double &Bel(int r, int c) { return sl.B[(size_t)c * PdBigSlab::LD + r]; }

    __device__ void push_rows(int r)
    {
        const bool ua = P.patt == 'U' || P.patt == 'B', ul = P.patt == 'L' || P.patt == 'B';
        const int na = ua ? pd_prefix(At, r, i) : 0;
        const int nb = ul ? pd_prefix(A, r, i) : 0;
        const int lane = threadIdx.x;
        for (int base = 0; base < na + nb; base += PD_WAVE) {
            const int q = base + lane;
            const bool have = q < na + nb;
            const int c = have ? (q < na ? At.idx[At.ptr[r] + q] : A.idx[A.ptr[r] + q - na]) : -1;
            sl.cand[lane] = c;
            pd_sync();
            bool isnew = have;
            for (int t = 0; t < lane && isnew; t++)
                if (sl.cand[t] == c) isnew = false;
            const int lim = cnt < PdBigSlab::LIST ? cnt : PdBigSlab::LIST;
            for (int t = 0; t < lim && isnew; t++)
                if (sl.list[t] == c) isnew = false;
            const unsigned long long m = pd_ballot(isnew);
            const int pos = cnt + pd_below(m);
            if (isnew && pos < PdBigSlab::LIST) sl.list[pos] = c;
            cnt += __popcll(m);
            if (cnt > PdBigSlab::LIST) ovf = true;
            pd_sync();
            if (ovf) return;
        }
    }

    __device__ void expand_snapshot()
    {
        const int len = cnt;
        for (int t = len - 1; t >= 0 && !ovf; t--) push_rows(sl.list[t]);
    }

    __device__ int rank_of(int c, int len) const
    {
        for (int t = 0; t < len; t++)
            if (sl.inds[t] == c) return t;
        return -1;
    }

    __device__ void solve(int len)
    {
#pragma clang fp contract(off)
        const int lane = threadIdx.x;
        for (int r = lane; r < len; r += PD_WAVE) {
            for (int c = 0; c < len; c++) Bel(r, c) = 0.0;
            Bel(r, r) = P.diag_floor;
            const int ar = sl.inds[r];
            for (int kk = A.ptr[ar]; kk < A.ptr[ar + 1] && A.idx[kk] < i; kk++) {
                const int k2 = rank_of(A.idx[kk], len);
                if (k2 >= 0) Bel(r, k2) = A.val[kk];
            }
            if (Bel(r, r) == 0.0) Bel(r, r) = P.diag_floor;
            sl.V[r] = 0.0;
            sl.W[r] = 0.0;
        }
        pd_sync();
        for (int q = At.ptr[i] + lane; q < At.ptr[i + 1]; q += PD_WAVE) {
            const int c = At.idx[q];
            if (c < i) {
                const int k2 = rank_of(c, len);
                if (k2 >= 0) sl.V[k2] = -At.val[q];
            }
        }
        for (int q = A.ptr[i] + lane; q < A.ptr[i + 1]; q += PD_WAVE) {
            const int c = A.idx[q];
            if (c < i) {
                const int k2 = rank_of(c, len);
                if (k2 >= 0) sl.W[k2] = -A.val[q];
            }
        }
        pd_sync();
        for (int j = 0; j < len; j++) {
            const double ajj = Bel(j, j);
            // this lane's rows in increasing order, then the wave: first index of the max
            double key = -1.0;
            int p = INT_MAX;
            for (int r = lane; r < len; r += PD_WAVE) {
                if (r < j) continue;
                const double a = fabs(Bel(r, j));
                const double k = (a == a) ? a : -1.0;
                if (k > key) {
                    key = k;
                    p = r;
                }
            }
            for (int off = 32; off >= 1; off >>= 1) {
                const double ko = __shfl_xor(key, off, PD_WAVE);
                const int po = __shfl_xor(p, off, PD_WAVE);
                if (ko > key || (ko == key && po < p)) {
                    key = ko;
                    p = po;
                }
            }
            if (!(fabs(ajj) == fabs(ajj))) p = j;
            else if (!(key > fabs(ajj))) p = j;
            if (lane == 0) sl.piv[j] = p;
            if (p != j)
                for (int c = lane; c < len; c += PD_WAVE) {
                    const double a = Bel(j, c), b = Bel(p, c);
                    Bel(j, c) = b;
                    Bel(p, c) = a;
                }
            pd_sync();
            const double d = Bel(j, j);
            if (d != 0.0)
                for (int r = lane; r < len; r += PD_WAVE) {
                    if (r <= j) continue;
                    const double l = Bel(r, j) / d;
                    Bel(r, j) = l;
                    if (l != 0.0)
                        for (int c = j + 1; c < len; c++) Bel(r, c) = Bel(r, c) - l * Bel(j, c);
                }
            pd_sync();
        }
        // B x = V: the swaps in order, then L and U column by column
        if (lane == 0)
            for (int j = 0; j < len; j++) {
                const int pj = sl.piv[j];
                if (pj != j) {
                    const double t = sl.V[j];
                    sl.V[j] = sl.V[pj];
                    sl.V[pj] = t;
                }
            }
        pd_sync();
        for (int c = 0; c < len; c++) {
            const double bc = sl.V[c];
            for (int r = lane; r < len; r += PD_WAVE)
                if (r > c) sl.V[r] = sl.V[r] - Bel(r, c) * bc;
            pd_sync();
        }
        for (int c = len - 1; c >= 0; c--) {
            if (lane == 0) sl.V[c] = sl.V[c] / Bel(c, c);
            pd_sync();
            const double bc = sl.V[c];
            for (int r = lane; r < c; r += PD_WAVE) sl.V[r] = sl.V[r] - Bel(r, c) * bc;
            pd_sync();
        }
        // B^T x = W: U^T, L^T, the swaps in reverse
        for (int c = 0; c < len; c++) {
            if (lane == 0) sl.W[c] = sl.W[c] / Bel(c, c);
            pd_sync();
            const double bc = sl.W[c];
            for (int r = lane; r < len; r += PD_WAVE)
                if (r > c) sl.W[r] = sl.W[r] - Bel(c, r) * bc;
            pd_sync();
        }
        for (int c = len - 1; c >= 0; c--) {
            const double bc = sl.W[c];
            for (int r = lane; r < c; r += PD_WAVE) sl.W[r] = sl.W[r] - Bel(c, r) * bc;
            pd_sync();
        }
        if (lane == 0)
            for (int j = len - 1; j >= 0; j--) {
                const int pj = sl.piv[j];
                if (pj != j) {
                    const double t = sl.W[j];
                    sl.W[j] = sl.W[pj];
                    sl.W[pj] = t;
                }
            }
        pd_sync();
    }

    __device__ double norm(const double *x, int len) const
    {
#pragma clang fp contract(off)
        double nrm = 1.0;
        if (P.filttype == '1') {
            for (int j = 0; j < len; j++) nrm += fabs(x[j]);
            nrm /= (len + 1);
        } else if (P.filttype == '2') {
            for (int j = 0; j < len; j++) nrm += x[j] * x[j];
            nrm = sqrt(nrm);
        } else {
            for (int j = 0; j < len; j++) nrm = (nrm < fabs(x[j])) ? fabs(x[j]) : nrm;
        }
        return nrm;
    }

    // survivors (flags fl) sorted by index into (si, sv), the unit diagonal i appended
    __device__ int store_sorted(const int *fl, const double *x, int len, int *si, double *sv)
    {
        const int lane = threadIdx.x;
        int c = 0;
        for (int r = 0; r < len; r++) c += fl[r];
        for (int r = lane; r < len; r += PD_WAVE) {
            if (!fl[r]) continue;
            const int me = sl.inds[r];
            int pos = 0;
            for (int t = 0; t < len; t++)
                if (fl[t] && sl.inds[t] < me) pos++;
            si[pos] = me;
            sv[pos] = x[r];
        }
        if (lane == 0) {
            si[c] = i;
            sv[c] = 1.0;
        }
        pd_sync();
        return c + 1;
    }

    __device__ void run(const PdOut &O, size_t slot)
    {
#pragma clang fp contract(off)
        const int lane = threadIdx.x;
        cnt = 0;
        push_rows(i);
        for (int k = 2; k <= P.prefill && !ovf; k++) expand_snapshot();
        double ng = 0.0, nh = 0.0;
        int tlen = 0;
        for (int k = 1; k <= P.lfill && !ovf; k++) {
            ng = 0.0;
            int len = cnt;
            tlen = len;
            if (len == 0) break;
            if (P.warpsize != -1 && len >= P.warpsize) cnt = P.warpsize - 1;
            len = cnt;
            tlen = len;
            if (len > PD_BIG) {
                ovf = true;
                break;
            }
            for (int r = lane; r < len; r += PD_WAVE) sl.inds[r] = sl.list[r];
            pd_sync();
            solve(len);
            ng = norm(sl.V, len);
            nh = norm(sl.W, len);
            for (int r = lane; r < len; r += PD_WAVE) {
                sl.GS[r] = fabs(sl.V[r]) >= P.ftol * ng;
                sl.HS[r] = fabs(sl.W[r]) >= P.ftol * nh;
            }
            pd_sync();
            // the set again: g's survivors, then h's that are not among them
            int c = 0, c2 = 0, cu = 0;
            for (int r = 0; r < len; r++) {
                c += sl.GS[r];
                c2 += sl.HS[r];
                cu += sl.HS[r] && !sl.GS[r];
            }
            int ng_pos = 0, nu_pos = c;
            for (int base = 0; base < len; base += PD_WAVE) {
                const int r = base + lane;
                const bool gs = r < len && sl.GS[r];
                const bool hu = r < len && sl.HS[r] && !sl.GS[r];
                const unsigned long long mg = pd_ballot(gs), mu = pd_ballot(hu);
                if (gs) sl.list[ng_pos + pd_below(mg)] = sl.inds[r];
                if (hu) sl.list[nu_pos + pd_below(mu)] = sl.inds[r];
                ng_pos += __popcll(mg);
                nu_pos += __popcll(mu);
            }
            cnt = c + cu;
            pd_sync();
            if ((c == 0 && c2 == 0) || c2 == P.warpsize) break;
            if (k == P.lfill) break;
            const int len_new = cnt;
            for (int j = 0; j < len_new && !ovf; j++) push_rows(sl.inds[j]);
            for (int m2 = 2; m2 <= P.postfill && !ovf; m2++) expand_snapshot();
            if (cnt == len_new) break;
        }
        if (ovf) {
            if (lane == 0) O.flag[i] = 1;
            return;
        }
        for (int r = lane; r < tlen; r += PD_WAVE) {
            sl.GS[r] = fabs(sl.V[r]) >= P.ftol * ng;
            sl.HS[r] = fabs(sl.W[r]) >= P.ftol * nh;
        }
        pd_sync();
        const int gn = store_sorted(sl.GS, sl.V, tlen, sl.gsi, sl.gsv);
        const int hn = store_sorted(sl.HS, sl.W, tlen, sl.hsi, sl.hsv);
        const size_t o = slot * O.out;
        for (int t = lane; t < gn; t += PD_WAVE) {
            O.gidx[o + t] = sl.gsi[t];
            O.gval[o + t] = sl.gsv[t];
        }
        for (int t = lane; t < hn; t += PD_WAVE) {
            O.hidx[o + t] = sl.hsi[t];
            O.hval[o + t] = sl.hsv[t];
        }
        double s = 0.0;
        for (int q = 0; q < hn; q++) {
            const int hj = sl.hsi[q];
            const double hv = sl.hsv[q];
            for (int b = A.ptr[hj]; b < A.ptr[hj + 1]; b += PD_WAVE) {
                const int kk = b + lane;
                double prod = 0.0;
                bool hit = false;
                if (kk < A.ptr[hj + 1]) {
                    const int col = A.idx[kk];
                    int lo = 0, hi = gn - 1;
                    while (lo <= hi) {
                        const int mid = (lo + hi) >> 1;
                        const int gm = sl.gsi[mid];
                        if (gm == col) {
                            hit = true;
                            prod = hv * A.val[kk] * sl.gsv[mid];
                            break;
                        }
                        if (gm < col) lo = mid + 1;
                        else hi = mid - 1;
                    }
                }
                const unsigned long long mh2 = pd_ballot(hit);
                for (int t = 0; t < PD_WAVE; t++) {
                    if (!((mh2 >> t) & 1ull)) continue;
                    s += pd_shfl(prod, t);
                }
            }
        }
        if (lane == 0) {
            O.gcnt[i] = gn;
            O.hcnt[i] = hn;
            O.dinv[i] = 1. / s;
            O.flag[i] = 0;
        }
    }
};

__global__ __launch_bounds__(PD_WAVE) void k_pdfiim_rows_big(int nrows, const int *__restrict__ rows, PdParams P,
                                                             PdCsr A, PdCsr At, PdOut O, PdBigSlab *slabs)
{
    PdBigSlab &sl = slabs[blockIdx.x];
    for (int q = blockIdx.x; q < nrows; q += gridDim.x) {
        PdRowBig row(sl, P, A, At, rows[q]);
        row.run(O, (size_t)q);
        __syncthreads();
    }
}

}  // namespace cm
