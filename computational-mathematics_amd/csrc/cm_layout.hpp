// Host-side layout of the engine's matrix formats: the constants that fix
// them and the index computations the engine uploads once per matrix --
// CSR validation, the staged SpMV's row blocks, the slot-major and diagonal
// indexes, the block products' row bins, triangular level schedules, the
// transpose and the symmetric expansion.  Plain C++ with no HIP, so the
// engine and a CPU harness built with the address and undefined-behaviour
// sanitizers (check/layout_check.cpp) run the same code.
#pragma once
#include <algorithm>
#include <climits>
#include <cstdint>
#include <utility>
#include <vector>

namespace cm {

constexpr int BS = 256;       // threads per workgroup (4 waves of 64)
constexpr int TILE = 2048;    // default nonzeros staged per SpMV workgroup (tunable: 1024, 2048, 4096)
constexpr int ROWCAP = 1024;  // rows per SpMV workgroup
constexpr int WAVE = 64;
constexpr int SLOT_ROWS = BS;     // rows per slot-major block
constexpr int SLOT_MAXLEN = 64;   // longest row the slot-major format takes
constexpr int DIA_MAX = 32;       // diagonals stored densely, at most
constexpr int DIA_ALIGN = 8;      // diagonal stride: n rounded up to this
constexpr double DIA_FILL = 0.6;  // a diagonal is stored densely when at least this share of rows has it
constexpr int SHORT_MAX = 64;     // block products: longer rows go to the wave-per-row bin
// Column tiles of the long remainders of the diagonal block products: about
// one tile per LONG_TILE_COLS columns (2 MB of x rows at 16 FP32 columns),
// at most LONG_TILES_MAX, and at least LONG_SEG_MIN remainder entries per
// (tile, row) segment on average.
constexpr int LONG_TILE_COLS = 32768, LONG_TILES_MAX = 32, LONG_SEG_MIN = 8;

// Why a CSR matrix would make a kernel read outside its arrays, or null.
// The row-pointer half of csr_problem (the column range is checked apart,
// on the device, when the columns are not read on the host).
inline const char *csr_ptr_problem(int n, const int *ptr, const int *col)
{
    if (n < 0 || !ptr) return "null CSR row pointer or negative size";
    if (ptr[0] != 0) return "CSR row pointer must start at 0";
    for (int i = 0; i < n; i++)
        if (ptr[i + 1] < ptr[i]) return "CSR row pointer not monotone";
    if (ptr[n] > 0 && !col) return "null CSR column array";
    return nullptr;
}

constexpr const char *CSR_COL_RANGE = "CSR column index out of range";

inline const char *csr_problem(int n, const int *ptr, const int *col)
{
    if (const char *why = csr_ptr_problem(n, ptr, col)) return why;
    for (long long k = 0; k < ptr[n]; k++)
        if (col[k] < 0 || col[k] >= n) return CSR_COL_RANGE;
    return nullptr;
}

// Row blocks for the adaptive SpMV: consecutive rows with at most `tile`
// nonzeros and ROWCAP rows, or a single longer row.  Returns the block
// boundaries (first row of every block, then n).
inline std::vector<int> make_row_blocks(int n, const int *ptr, int tile = TILE)
{
    std::vector<int> b;
    b.push_back(0);
    int rows = 0;
    long long nz = 0;
    for (int r = 0; r < n; r++) {
        const int len = ptr[r + 1] - ptr[r];
        if (rows > 0 && (nz + len > tile || rows >= ROWCAP)) {
            b.push_back(r);
            rows = 0;
            nz = 0;
        }
        rows++;
        nz += len;
        if (len > tile) {
            b.push_back(r + 1);
            rows = 0;
            nz = 0;
        }
    }
    if (rows > 0) b.push_back(n);
    return b;
}

// The SpMV kernels' row-block table: the block row starts, then the offsets
// of the blocks' first nonzeros (ptr at those rows); 2 (nblk + 1) entries.
inline std::vector<int> block_table(int n, const int *ptr, int *nblk, int tile = TILE)
{
    std::vector<int> b = make_row_blocks(n, ptr, tile);
    const size_t nb1 = b.size();
    *nblk = (int)nb1 - 1;
    b.resize(2 * nb1);
    for (size_t i = 0; i < nb1; i++) b[nb1 + i] = ptr[b[i]];
    return b;
}

// Rows with at most short_max nonzeros, and the longer ones.
inline void row_bins(int n, const int *ptr, int short_max, std::vector<int> &sh, std::vector<int> &lo)
{
    sh.clear();
    lo.clear();
    for (int i = 0; i < n; i++) (ptr[i + 1] - ptr[i] > short_max ? lo : sh).push_back(i);
}

// CSR of the transpose.
inline void host_transpose(int n, const int *ptr, const int *col, const double *val, std::vector<int> &tp,
                           std::vector<int> &tc, std::vector<double> &tv)
{
    tp.assign(n + 1, 0);
    for (int k = 0; k < ptr[n]; k++) tp[col[k] + 1]++;
    for (int i = 0; i < n; i++) tp[i + 1] += tp[i];
    tc.resize(ptr[n]);
    tv.resize(ptr[n]);
    std::vector<int> fill(tp.begin(), tp.end() - 1);
    for (int i = 0; i < n; i++)
        for (int k = ptr[i]; k < ptr[i + 1]; k++) {
            int q = fill[col[k]]++;
            tc[q] = i;
            tv[q] = val[k];
        }
}

// Full CSR of a symmetric matrix given its upper triangle (entries below the
// diagonal, if any, are ignored as the SYMMETRIC/UPPER descriptor does).
inline void host_sym_expand(int n, const int *ptr, const int *col, const double *val, std::vector<int> &fp,
                            std::vector<int> &fc, std::vector<double> &fv)
{
    fp.assign(n + 1, 0);
    for (int i = 0; i < n; i++)
        for (int k = ptr[i]; k < ptr[i + 1]; k++) {
            const int j = col[k];
            if (j < i) continue;
            fp[i + 1]++;
            if (j != i) fp[j + 1]++;
        }
    for (int i = 0; i < n; i++) fp[i + 1] += fp[i];
    fc.resize(fp[n]);
    fv.resize(fp[n]);
    std::vector<int> fill(fp.begin(), fp.end() - 1);
    // lower part first (column order ascending because rows are scanned in
    // order), then the stored upper part of the row
    for (int i = 0; i < n; i++)
        for (int k = ptr[i]; k < ptr[i + 1]; k++) {
            const int j = col[k];
            if (j <= i) continue;
            int q = fill[j]++;
            fc[q] = i;
            fv[q] = val[k];
        }
    for (int i = 0; i < n; i++)
        for (int k = ptr[i]; k < ptr[i + 1]; k++) {
            const int j = col[k];
            if (j < i) continue;
            int q = fill[i]++;
            fc[q] = j;
            fv[q] = val[k];
        }
}

// Dependency levels of a triangular factor: lower -> a row reads only
// smaller columns (rows in increasing order); upper -> larger columns.
// Entries on the other side of the diagonal are ignored.  rows[loff[l] ..
// loff[l + 1]) are level l's rows.  The declared triangle's entries
// (diagonal included) are also stored in that row order, tptr / tcol / tsrc
// (tsrc: the entry's CSR position), so a level's rows read their entries
// from one contiguous run.
struct LevelSchedule {
    int nlev = 0;
    std::vector<int> loff, rows, tptr, tcol, tsrc;
};

// (with_triangle false: loff and rows only; the engine gathers the
// triangle's entries on the device)
inline LevelSchedule level_schedule(int n, const int *ptr, const int *col, bool lower, bool with_triangle = true)
{
    LevelSchedule L;
    std::vector<int> lev(n, 0);
    for (int q = 0; q < n; q++) {
        const int i = lower ? q : n - 1 - q;
        int l = 0;
        for (int k = ptr[i]; k < ptr[i + 1]; k++) {
            const int j = col[k];
            if ((lower && j < i) || (!lower && j > i)) l = std::max(l, lev[j] + 1);
        }
        lev[i] = l;
        L.nlev = std::max(L.nlev, l + 1);
    }
    L.loff.assign(L.nlev + 1, 0);
    for (int i = 0; i < n; i++) L.loff[lev[i] + 1]++;
    for (int l = 0; l < L.nlev; l++) L.loff[l + 1] += L.loff[l];
    L.rows.assign(n, 0);
    std::vector<int> fill(L.loff.begin(), L.loff.end() - 1);
    for (int i = 0; i < n; i++) L.rows[fill[lev[i]]++] = i;
    if (!with_triangle) return L;
    L.tptr.assign(n + 1, 0);
    L.tcol.clear();
    L.tsrc.clear();
    for (int q = 0; q < n; q++) {
        const int i = L.rows[q];
        for (int k = ptr[i]; k < ptr[i + 1]; k++) {
            const int j = col[k];
            if (lower ? j <= i : j >= i) {
                L.tcol.push_back(j);
                L.tsrc.push_back(k);
            }
        }
        L.tptr[q + 1] = (int)L.tcol.size();
    }
    return L;
}

// Slot-major index (spmv_slot): blocks of SLOT_ROWS rows, rows ordered by
// length inside a block (longest first, ties in row order), nonzeros slot by
// slot.  hdr holds per block (first row, rows, first nonzero, first slot
// offset, longest row); soff the per-slot offsets; lrow per row (length << 16
// | position in the block's order); col / src the columns and CSR positions in
// slot order; d16 / cbase the 16-bit offsets from the block's smallest
// column when every block's columns span less than 65536.
struct SlotLayout {
    int nb = 0;
    bool fits16 = true;
    std::vector<int> hdr, soff, lrow, col, src, cbase;
    std::vector<uint16_t> d16;
};

// false when a row is longer than SLOT_MAXLEN (the format does not apply)
inline bool slot_layout(int n, const int *ptr, const int *col, SlotLayout &S)
{
    for (int i = 0; i < n; i++)
        if (ptr[i + 1] - ptr[i] > SLOT_MAXLEN) return false;
    const long long nnz = ptr[n];
    const int nb = (n + SLOT_ROWS - 1) / SLOT_ROWS;
    S.nb = nb;
    S.fits16 = true;
    S.hdr.assign((size_t)nb * 5, 0);
    S.soff.clear();
    S.lrow.assign(n, 0);
    S.col.assign(nnz, 0);
    S.src.assign(nnz, 0);
    S.d16.assign(nnz, 0);
    S.cbase.assign(nb, 0);
    long long pos = 0;
    std::vector<int> ord(SLOT_ROWS);
    for (int b = 0; b < nb; b++) {
        const int r0 = b * SLOT_ROWS, nr = std::min(SLOT_ROWS, n - r0);
        auto len = [&](int r) { return ptr[r + 1] - ptr[r]; };
        for (int i = 0; i < nr; i++) ord[i] = i;
        std::stable_sort(ord.begin(), ord.begin() + nr, [&](int a, int c) { return len(r0 + a) > len(r0 + c); });
        const int L = nr ? len(r0 + ord[0]) : 0;
        int lo = INT32_MAX, hi = -1;
        for (long long k = ptr[r0]; k < ptr[r0 + nr]; k++) {
            lo = std::min(lo, col[k]);
            hi = std::max(hi, col[k]);
        }
        if (hi >= 0 && hi - lo > 65535) S.fits16 = false;
        S.cbase[b] = hi >= 0 ? lo : 0;
        S.hdr[5 * (size_t)b + 0] = r0;
        S.hdr[5 * (size_t)b + 1] = nr;
        S.hdr[5 * (size_t)b + 2] = (int)pos;
        S.hdr[5 * (size_t)b + 3] = (int)S.soff.size();
        S.hdr[5 * (size_t)b + 4] = L;
        int off = 0;
        for (int j = 0; j < L; j++) {
            S.soff.push_back(off);
            for (int i = 0; i < nr; i++) {
                const int r = r0 + ord[i];
                if (len(r) <= j) break;  // rows are in decreasing length
                const long long k = ptr[r] + j;
                S.col[pos] = col[k];
                S.src[pos] = (int)k;
                S.d16[pos] = (uint16_t)(col[k] - S.cbase[b]);
                pos++;
                off++;
            }
        }
        S.soff.push_back(off);
        for (int i = 0; i < nr; i++) S.lrow[r0 + i] = (len(r0 + ord[i]) << 16) | ord[i];
    }
    return true;
}

// Diagonal index (spmv_dia): the diagonals held by at least DIA_FILL of the
// rows (the DIA_MAX most populated, in offset order) as dense vectors of
// stride ldn (dsrc: the CSR position of every diagonal element, -1 where
// absent or outside the matrix), the other entries as a CSR remainder (rptr,
// rcol, rsrc).  Rows whose remainder is longer than SHORT_MAX (the block
// products' long rows: lpos[row] = their index, -1 for the others) also get
// their remainders cut into ntile column tiles of tw columns, stored as a
// tile-major CSR of (tile, row) segments (lptr, lcol, lsrc; segment t * nlong
// + q holds long row q's entries with columns in [t tw, (t + 1) tw), in row
// order).
struct DiaLayout {
    int ldn = 0, nd = 0, nlong = 0, ntile = 1, tw = 0;
    long long nseg = 0;
    std::vector<int> off, dsrc, rptr, rcol, rsrc, lpos, lptr, lcol, lsrc;
};

// The stored diagonals from (entry count, offset) pairs: those held by at
// least DIA_FILL of the n rows, the DIA_MAX most populated (ties: the
// smaller offset), in offset order.
inline std::vector<int> dia_select(int n, std::vector<std::pair<long long, int>> cand)
{
    cand.erase(std::remove_if(cand.begin(), cand.end(),
                              [n](const std::pair<long long, int> &c) { return !((double)c.first >= DIA_FILL * n); }),
               cand.end());
    std::sort(cand.begin(), cand.end(), [](const auto &a, const auto &b) {
        return a.first != b.first ? a.first > b.first : a.second < b.second;
    });
    if ((int)cand.size() > DIA_MAX) cand.resize(DIA_MAX);
    std::vector<int> off;
    for (const auto &c : cand) off.push_back(c.second);
    std::sort(off.begin(), off.end());
    return off;
}

// false when no diagonal qualifies (the format does not apply)
inline bool dia_layout(int n, const int *ptr, const int *col, DiaLayout &D)
{
    if (n <= 0) return false;
    // entries per diagonal, indexed by offset + n - 1 (offsets span 1 - n .. n - 1)
    std::vector<int> cnt(2 * (size_t)n - 1, 0);
    for (int i = 0; i < n; i++)
        for (int k = ptr[i]; k < ptr[i + 1]; k++) cnt[(size_t)(col[k] - i + n - 1)]++;
    std::vector<std::pair<long long, int>> cand;
    for (size_t o = 0; o < cnt.size(); o++)
        if (cnt[o] > 0) cand.push_back({cnt[o], (int)((long long)o - (n - 1))});
    D.off = dia_select(n, std::move(cand));
    if (D.off.empty()) return false;
    D.nd = (int)D.off.size();
    D.ldn = (n + DIA_ALIGN - 1) / DIA_ALIGN * DIA_ALIGN;
    // diagonal index by offset (reusing the count array): -1 for the others
    std::fill(cnt.begin(), cnt.end(), -1);
    for (int d = 0; d < D.nd; d++) cnt[(size_t)(D.off[d] + n - 1)] = d;
    D.dsrc.assign((size_t)D.nd * D.ldn, -1);
    D.rptr.assign(n + 1, 0);
    D.rcol.clear();
    D.rsrc.clear();
    for (int i = 0; i < n; i++) {
        for (int k = ptr[i]; k < ptr[i + 1]; k++) {
            const int d = cnt[(size_t)(col[k] - i + n - 1)];
            if (d >= 0) {
                D.dsrc[(size_t)d * D.ldn + i] = k;
            } else {
                D.rcol.push_back(col[k]);
                D.rsrc.push_back(k);
            }
        }
        D.rptr[i + 1] = (int)D.rcol.size();
    }
    // long remainders (block products), as a tile-major CSR of (column
    // tile, row) segments (counting sort by tile)
    D.lpos.assign(n, -1);
    std::vector<int> lrow;
    long long nlong_rem = 0;
    for (int i = 0; i < n; i++)
        if (D.rptr[i + 1] - D.rptr[i] > SHORT_MAX) {
            D.lpos[i] = (int)lrow.size();
            lrow.push_back(i);
            nlong_rem += D.rptr[i + 1] - D.rptr[i];
        }
    D.nlong = (int)lrow.size();
    int ntile = std::min(LONG_TILES_MAX, std::max(1, (n + LONG_TILE_COLS - 1) / LONG_TILE_COLS));
    // at least LONG_SEG_MIN entries per segment on average: bounds the
    // segment sums (written and read back by every block product, and held
    // in the block workspace) by the long remainders' own size
    if (D.nlong)
        ntile = (int)std::min<long long>(ntile, std::max<long long>(1, nlong_rem / ((long long)LONG_SEG_MIN * D.nlong)));
    if ((long long)ntile * D.nlong > INT_MAX) ntile = 1;  // segment ids are int
    D.ntile = ntile;
    D.tw = (n + ntile - 1) / ntile;
    D.nseg = (long long)ntile * D.nlong;
    D.lptr.assign((size_t)D.nseg + 1, 0);
    for (int q = 0; q < D.nlong; q++)
        for (int k = D.rptr[lrow[q]]; k < D.rptr[lrow[q] + 1]; k++)
            D.lptr[(size_t)(D.rcol[k] / D.tw) * D.nlong + q + 1]++;
    for (long long g = 0; g < D.nseg; g++) D.lptr[g + 1] += D.lptr[g];
    D.lcol.assign(D.lptr[D.nseg], 0);
    D.lsrc.assign(D.lptr[D.nseg], 0);
    std::vector<int> fill(D.lptr.begin(), D.lptr.end() - 1);
    for (int q = 0; q < D.nlong; q++)
        for (int k = D.rptr[lrow[q]]; k < D.rptr[lrow[q] + 1]; k++) {
            const int pos = fill[(size_t)(D.rcol[k] / D.tw) * D.nlong + q]++;
            D.lcol[pos] = D.rcol[k];
            D.lsrc[pos] = D.rsrc[k];
        }
    return true;
}

}  // namespace cm
