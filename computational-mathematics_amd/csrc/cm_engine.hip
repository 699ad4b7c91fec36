// Host driver and C ABI of the BiCGSTAB solve engine (one GPU, gfx950).
//
// A solve is a device-resident state machine: the host uploads the system
// once, enqueues the setup kernel and then chunks of iterations on the
// system's own stream, and reads the solve state back once per chunk (at
// most two chunks outstanding), so the GPU never waits on the host inside a
// chunk.  Every kernel returns immediately once the state says done, so
// enqueuing past the exit is harmless and the iteration at which the loop
// stops is the reference's, decided on the device.
//
// Per BiCGSTAB pass (mode E / AFIIM / PDFIIM), 12 launches:
//   tm = D^{-1}(H p)   pc = G tm   v = A pc (+ r0.v)   [alpha]   s = r - alpha v
//   tm = D^{-1}(H s)   sc = G tm   t = A sc (+ t.s, t.t)   [omega]
//   x, r update (+ ||pc||, ||x||, ||r||, r0.r)   [exit tests, beta]   p update
// ([..] = one-workgroup finalize of the preceding reduction)
// (AFIIM/src/bicgstab.cpp:168-275, PDFIIM/src/bicgstab.cpp:186-293,
//  ACIFIM/src/bicgstab.cpp:197-304).
//
// Two solve paths share the Krylov core (Krylov<T, VA, VF>: vector, A-value
// and factor-value types):
//  * FP64 (double everywhere) -- the drop-in replacement of the reference;
//  * mixed precision -- FP64 iterative refinement (true residual b - A x,
//    solution and norm in FP64, with the FP64 A) around an inner BiCGSTAB on
//    FP32 A, FP32 or FP16 (per-row scaled) factors and FP32 vectors.  The
//    low-precision matrices are cast on the device from the resident FP64
//    ones and share their index arrays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "cm_kernels.hpp"
#include "cm_solve.h"
#include "cm_batch.hpp"
#include "cm_block.hpp"
#include "cm_trsv.hpp"
#include "cm_pdfiim.hpp"

namespace cm {

static thread_local std::string g_last_error;

static int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

#define CM_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(CM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));     \
    } while (0)

#define CM_TRY(expr)                  \
    do {                              \
        int rc_ = (expr);             \
        if (rc_ != CM_OK) return rc_; \
    } while (0)

// ------------------------------------------------------ device-side casts

// dst = (V) src
template <class V>
__global__ __launch_bounds__(BS) void k_cast(long long cnt, const double *__restrict__ src, V *__restrict__ dst)
{
    for (long long i = (long long)blockIdx.x * BS + threadIdx.x; i < cnt; i += (long long)gridDim.x * BS)
        dst[i] = (V)src[i];
}

// Row-scaled FP16 copy: scale[i] = max |row i| (1 where the row is zero),
// dst = src / scale[i] rounded to FP16 (so |dst| <= 1).
__global__ __launch_bounds__(BS) void k_half_rows(int n, const int *__restrict__ ptr, const double *__restrict__ src,
                                                  __half *__restrict__ dst, float *__restrict__ scale)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i >= n) return;
    double mx = 0.0;
    for (int k = ptr[i]; k < ptr[i + 1]; k++) mx = fmax(mx, fabs(src[k]));
    const float sc = mx > 0.0 ? (float)mx : 1.0f;
    scale[i] = sc;
    for (int k = ptr[i]; k < ptr[i + 1]; k++) dst[k] = __float2half_rn((float)(src[k] / (double)sc));
}

// dst[k] = src[idx[k]] (0 where idx[k] < 0): a matrix's values in the
// slot-major or diagonal order
template <class V>
__global__ __launch_bounds__(BS) void k_gather_vals(long long cnt, const int *__restrict__ idx, const V *__restrict__ src,
                                                    V *__restrict__ dst)
{
    for (long long i = (long long)blockIdx.x * BS + threadIdx.x; i < cnt; i += (long long)gridDim.x * BS) {
        const int j = idx[i];
        dst[i] = j >= 0 ? src[j] : (V)0.0f;
    }
}

// z = a .* d (mode I of the ACIFIM apply)
template <class T>
__global__ __launch_bounds__(BS) void k_vmul(int n, const T *__restrict__ a, const T *__restrict__ d,
                                             T *__restrict__ z, const SolveState *st)
{
    if (st && st->done) return;
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) z[i] = a[i] * d[i];
}

// Sets *bad to 1 when a stored column lies outside [0, n) (the upload's
// column check, on the device).
__global__ __launch_bounds__(BS) void k_col_range(long long nnz, const int *__restrict__ col, int n, int *bad)
{
    bool out = false;
    for (long long k = (long long)blockIdx.x * BS + threadIdx.x; k < nnz; k += (long long)gridDim.x * BS)
        out |= (unsigned)col[k] >= (unsigned)n;
    if (out) *bad = 1;
}

// ------------------------------------------------------------ host helpers

template <class T>
static int dev_alloc(T **d, size_t count)
{
    *d = nullptr;
    CM_HIP(hipMalloc((void **)d, std::max<size_t>(count, 1) * sizeof(T)));
    return CM_OK;
}

// Host-to-device copies of large arrays through pinned staging chunks: the
// calling thread and STAGE_HELPERS helper threads copy the (pageable) source
// into one chunk while the DMA engine moves the previous ones -- about twice
// the rate of a pageable hipMemcpyAsync, which stages through the runtime's
// own buffers with one thread.  One per process, made on first use and kept
// until the process ends; uploads from several threads take turns.  Returns
// once the source has been read; the device copies complete in the order of
// the stream passed (a chunk is refilled only after the event recorded
// behind its last copy, whichever upload that was).
class Stager {
  public:
    static constexpr size_t CHUNK = (size_t)16 << 20, MIN_BYTES = (size_t)4 << 20;
    static constexpr int NBUF = 4, STAGE_HELPERS = 3;
    static Stager &get()
    {
        static Stager *s = new Stager();  // never destroyed: its helpers wait until the process ends
        return *s;
    }
    int upload(void *dst, const void *src, size_t bytes, hipStream_t s)
    {
        std::lock_guard<std::mutex> lk(mu);
        if (!buf[0]) {
            for (auto &b : buf) CM_HIP(hipHostMalloc((void **)&b, CHUNK, hipHostMallocDefault));
            for (int h = 0; h < STAGE_HELPERS; h++) std::thread([this] { help(); }).detach();
        }
        for (size_t off = 0; off < bytes; next = (next + 1) % NBUF) {
            const int b = next;
            const size_t len = std::min(CHUNK, bytes - off);
            if (ev[b]) {  // the chunk's previous copy must have left it
                const hipError_t e = hipEventSynchronize(ev[b]);
                (void)hipEventDestroy(ev[b]);
                ev[b] = nullptr;
                if (e != hipSuccess) return fail(CM_ERR_HIP, "staged upload wait failed");
            }
            copy(buf[b], (const char *)src + off, len);
            CM_HIP(hipMemcpyAsync((char *)dst + off, buf[b], len, hipMemcpyHostToDevice, s));
            CM_HIP(hipEventCreateWithFlags(&ev[b], hipEventDisableTiming));  // on the stream's (current) device
            CM_HIP(hipEventRecord(ev[b], s));
            off += len;
        }
        return CM_OK;
    }

  private:
    std::mutex mu;  // one upload at a time
    char *buf[NBUF] = {};
    hipEvent_t ev[NBUF] = {};  // recorded behind each chunk's last copy
    int next = 0;              // the chunk filled next
    struct Job {
        char *d;
        const char *s;
        size_t n;
    };
    std::mutex jmu;
    std::condition_variable jcv, dcv;
    std::vector<Job> jobs;
    int pending = 0;

    // len bytes in STAGE_HELPERS + 1 parts, the first by the caller
    void copy(char *d, const char *src, size_t len)
    {
        const size_t part = ((len + STAGE_HELPERS) / (STAGE_HELPERS + 1) + 4095) / 4096 * 4096;
        {
            std::lock_guard<std::mutex> lk(jmu);
            for (size_t a = part; a < len; a += part) {
                jobs.push_back({d + a, src + a, std::min(part, len - a)});
                pending++;
            }
        }
        jcv.notify_all();
        std::memcpy(d, src, std::min(part, len));
        std::unique_lock<std::mutex> lk(jmu);
        dcv.wait(lk, [this] { return pending == 0; });
    }
    void help()
    {
        for (;;) {
            Job j;
            {
                std::unique_lock<std::mutex> lk(jmu);
                jcv.wait(lk, [this] { return !jobs.empty(); });
                j = jobs.back();
                jobs.pop_back();
            }
            std::memcpy(j.d, j.s, j.n);
            std::lock_guard<std::mutex> lk(jmu);
            if (--pending == 0) dcv.notify_all();
        }
    }
};

// count T from host memory h into device memory d (staged when large)
template <class T>
static int copy_h2d(T *d, const T *h, size_t count, hipStream_t s)
{
    const size_t bytes = count * sizeof(T);
    if (bytes >= Stager::MIN_BYTES) return Stager::get().upload(d, h, bytes, s);
    CM_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s));
    return CM_OK;
}

template <class T>
static int dev_upload(T **d, const T *h, size_t count, hipStream_t s)
{
    CM_TRY(dev_alloc(d, count));
    if (h && count) CM_TRY(copy_h2d(*d, h, count, s));
    return CM_OK;
}

// Matrix arrays: nnz values followed by CSR_PAD zeros (the 16-byte loads of
// spmv_vl may read past the last nonzero).
template <class T>
static int dev_upload_padded(T **d, const T *h, size_t count, hipStream_t s)
{
    CM_TRY(dev_alloc(d, count + CSR_PAD));
    if (h && count) CM_TRY(copy_h2d(*d, h, count, s));
    CM_HIP(hipMemsetAsync(*d + count, 0, CSR_PAD * sizeof(T), s));
    return CM_OK;
}

static inline int vec_grid(int n)
{
    int g = (n + BS - 1) / BS;
    return std::max(1, std::min(g, MAX_RED_BLOCKS));
}

static inline int long_grid(long long cnt)
{
    long long g = (cnt + BS - 1) / BS;
    return (int)std::max<long long>(1, std::min<long long>(g, 8192));
}

// Index arrays of a matrix's slot-major copy (spmv_slot), shared by its
// FP64 and low-precision copies; each copy holds its own values in that
// order, gathered through src (the CSR position of every slot-major nonzero).
struct SlotIndex {
    int n = 0, nb = 0;
    long long nnz = 0;
    size_t nsoff = 0;
    int *hdr = nullptr, *soff = nullptr, *lrow = nullptr, *col = nullptr, *cbase = nullptr, *src = nullptr;
    uint16_t *d16 = nullptr;  // null when a block's columns span 65536 or more
    size_t bytes() const
    {
        return ((size_t)nb * 5 + nsoff + (size_t)n + 2 * (size_t)nnz) * sizeof(int) +
               (d16 ? (size_t)nnz * sizeof(uint16_t) + (size_t)nb * sizeof(int) : 0);
    }
    ~SlotIndex()
    {
        (void)hipFree(hdr);
        (void)hipFree(soff);
        (void)hipFree(lrow);
        (void)hipFree(col);
        (void)hipFree(cbase);
        (void)hipFree(src);
        (void)hipFree(d16);
    }
};

// Index arrays of a matrix's diagonal copy (spmv_dia), shared like
// SlotIndex: the diagonal offsets, for every diagonal element the CSR
// position of its entry (-1: none, stored as 0), and the CSR remainder of
// the entries off those diagonals with their CSR positions.
struct DiaIndex {
    int n = 0, ldn = 0, nd = 0;  // ldn: diagonal stride, n rounded up to DIA_ALIGN
    long long nrem = 0;
    int *off = nullptr, *dsrc = nullptr, *rptr = nullptr, *rcol = nullptr, *rsrc = nullptr;
    // Rows whose remainder is longer than SHORT_MAX, for the block products:
    // lpos[row] = their index q (0..nlong-1), -1 for the other rows.  Their
    // remainders are cut into ntile column tiles of tw columns and stored as a
    // CSR of segments, tile-major: segment t * nlong + q holds row q's entries
    // with columns in [t * tw, (t + 1) * tw) (lptr / lcol, values gathered
    // through lsrc); the short-row kernel sums them, segment q as its row q.
    // Walking the segments in order keeps the x rows that an XCD's waves in
    // flight gather within about one tile, which its L2 holds.
    int nlong = 0, ntile = 1, tw = 0;
    long long nlrem = 0, nseg = 0;
    int *lpos = nullptr, *lptr = nullptr, *lcol = nullptr, *lsrc = nullptr;
    size_t bytes() const
    {
        return ((size_t)nd + (size_t)nd * ldn + (rptr ? (size_t)n + 1 : 0) + 2 * (size_t)nrem +
                (nlong ? (size_t)n + (size_t)nseg + 1 + 2 * (size_t)nlrem : 0)) *
               sizeof(int);
    }
    ~DiaIndex()
    {
        (void)hipFree(off);
        (void)hipFree(dsrc);
        (void)hipFree(rptr);
        (void)hipFree(rcol);
        (void)hipFree(rsrc);
        (void)hipFree(lpos);
        (void)hipFree(lptr);
        (void)hipFree(lcol);
        (void)hipFree(lsrc);
    }
};

// SpMV formats of a resident matrix copy: the staged row blocks with 32-bit
// or 16-bit column indices, the slot-major blocks with either, or the
// diagonals with a CSR remainder.
enum SpmvFormat { FMT_STAGED = 0, FMT_STAGED16 = 1, FMT_SLOT = 2, FMT_SLOT16 = 3, FMT_DIA = 4, FMT_DIA2 = 5 };

// Rows per thread of spmv_dia: FP16 values are read two rows at a time (one
// 4-byte load per diagonal); FP32 and FP64 one row per thread.  The second
// diagonal variant (FMT_DIA2) doubles the rows per thread and takes 9
// diagonals per chunk instead of 8: faster on matrices with many diagonals
// (the 17 of the 128^3 PDFIIM H factor, in two chunks), slower on few.
template <class V>
constexpr int dia_rpt() { return std::is_same<V, __half>::value ? 2 : 1; }
// The second diagonal variant is offered (to the autotuner, and picked by a
// one-shot call) from this many diagonals up: on fewer (a 7-point stencil's
// 7) it timed within noise of the first in isolation and 1.5 µs slower in
// the solve, so the autotuner's pick flipped from run to run.
constexpr int DIA2_MIN_ND = 12;

// A CSR matrix in HBM with its SpMV row blocks (and, for triangular factors
// used by the level-scheduled solves, its level schedule).  A low-precision
// copy shares ptr/col/blk/levels with the FP64 matrix it was cast from
// (owns_index = false).  FP16 values are stored divided by the row's largest
// magnitude, which rowscale holds.
template <class V>
struct DevCsr {
    int n = 0;
    long long nnz = 0;
    int nblk = 0;
    int tile = TILE;  // staged nonzeros per SpMV workgroup the row blocks were cut for
    bool owns_index = true;
    int *ptr = nullptr, *col = nullptr;
    int *blk = nullptr;  // SpMV row-block table: block row starts, then their first nonzeros
    uint16_t *d16 = nullptr;  // optional block-relative 16-bit columns (col - cbase[block])
    int *cbase = nullptr;
    bool use16 = false;       // low-precision copy: products through spmv_vl16
    V *val = nullptr;
    float *rowscale = nullptr;
    std::shared_ptr<SlotIndex> slots;  // slot-major index (shared by the copies), if built
    V *sval = nullptr;                 // this copy's values in slot-major order, if built
    bool use_slot = false;             // products through spmv_slot
    bool slot16 = false;               // ... with its 16-bit column offsets
    std::shared_ptr<DiaIndex> dia;     // diagonal index (shared by the copies), if built
    V *dval = nullptr, *rval = nullptr;  // this copy's diagonal and remainder values, if built
    V *lval = nullptr;                   // ... and its long-row remainder values (block products)
    bool use_dia = false;              // products through spmv_dia
    bool dia2 = false;                 // ... in its second variant (FMT_DIA2)
    bool block_dia = false;            // block products (SpMM) through spmm_dia
    int nlev = 0;
    bool lower = false;  // the level schedule's triangle (lower: rows read smaller columns)
    int *lrows = nullptr;
    std::vector<int> loff;
    int *dloff = nullptr;         // loff on the device (read by the run-of-levels kernel)
    std::vector<long long> lnnz;  // per level: the triangle's entries of its rows
    int lpr = 1;                  // lanes per row of the level kernel
    // the triangle's entries in level order (tptr indexed by level position;
    // tsrc: CSR positions; tval: this copy's values gathered through tsrc)
    int *tptr = nullptr, *tcol = nullptr, *tsrc = nullptr;
    long long ntri = 0;
    V *tval = nullptr;
    bool present() const { return ptr != nullptr; }
    // device bytes held (index arrays only when owned)
    size_t bytes() const
    {
        if (!present()) return 0;
        size_t b = (size_t)nnz * sizeof(V) + (rowscale ? (size_t)n * sizeof(float) : 0);
        if (sval) b += (size_t)nnz * sizeof(V);
        if (dval && dia) b += ((size_t)dia->nd * dia->ldn + (size_t)dia->nrem + (size_t)dia->nlrem) * sizeof(V);
        if (tval) b += (size_t)ntri * sizeof(V);
        if (owns_index) {
            if (tptr) b += ((size_t)n + 1 + 2 * (size_t)ntri) * sizeof(int);
            if (slots) b += slots->bytes();
            if (dia) b += dia->bytes();
            b += (size_t)(n + 1 + 2 * (nblk + 1)) * sizeof(int) + (size_t)nnz * sizeof(int) + (lrows ? (size_t)n * sizeof(int) : 0);
            if (dloff) b += (size_t)(nlev + 1) * sizeof(int);
            if (d16) b += (size_t)(nnz + 2 * CSR_PAD) * sizeof(uint16_t) + (size_t)nblk * sizeof(int);
        }
        return b;
    }
    void release()
    {
        if (owns_index) {
            (void)hipFree(ptr);
            (void)hipFree(col);
            (void)hipFree(blk);
            (void)hipFree(lrows);
            (void)hipFree(dloff);
            (void)hipFree(d16);
            (void)hipFree(cbase);
            (void)hipFree(tptr);
            (void)hipFree(tcol);
            (void)hipFree(tsrc);
        }
        (void)hipFree(tval);
        (void)hipFree(val);
        (void)hipFree(rowscale);
        (void)hipFree(sval);
        (void)hipFree(dval);
        (void)hipFree(rval);
        (void)hipFree(lval);
        ptr = col = blk = lrows = cbase = dloff = nullptr;
        tptr = tcol = tsrc = nullptr;
        ntri = 0;
        d16 = nullptr;
        val = sval = dval = rval = lval = tval = nullptr;
        rowscale = nullptr;
        slots.reset();
        dia.reset();
        use_slot = slot16 = use_dia = block_dia = false;
    }
    ~DevCsr() { release(); }
};

static int check_csr(int n, const int *ptr, const int *col)
{
    // The kernels index x by the stored columns and stage rows by ptr: reject
    // anything that would read outside the arrays.
    if (const char *why = csr_problem(n, ptr, col)) return fail(CM_ERR_ARG, why);
    return CM_OK;
}

// Uploads a CSR matrix.  bad: null to check the columns on the host (for
// callers that go on reading them there), else a device flag that the
// columns' range check sets, which the caller reads back before any kernel
// uses the matrix.
static int upload_csr(DevCsr<double> &m, int n, const int *ptr, const int *col, const double *val, hipStream_t s,
                      int *bad = nullptr)
{
    if (!bad) {
        CM_TRY(check_csr(n, ptr, col));
    } else if (const char *why = csr_ptr_problem(n, ptr, col)) {
        return fail(CM_ERR_ARG, why);
    }
    m.release();
    m.owns_index = true;
    m.n = n;
    m.nnz = ptr[n];
    // the staged row blocks are cut on another thread while the arrays upload
    std::vector<int> blk;
    int nblk = 0;
    struct Joined {
        std::thread t;
        ~Joined()
        {
            if (t.joinable()) t.join();
        }
    } cut{std::thread([&] { blk = block_table(n, ptr, &nblk); })};
    CM_TRY(dev_upload(&m.ptr, ptr, (size_t)n + 1, s));
    CM_TRY(dev_upload_padded(&m.col, col, (size_t)m.nnz, s));
    if (bad && m.nnz) k_col_range<<<long_grid(m.nnz), BS, 0, s>>>(m.nnz, m.col, n, bad);
    CM_TRY(dev_upload_padded(&m.val, val, (size_t)m.nnz, s));
    cut.t.join();
    m.nblk = nblk;
    CM_TRY(dev_upload(&m.blk, blk.data(), blk.size(), s));
    CM_HIP(hipStreamSynchronize(s));  // blk is a host temporary
    return CM_OK;
}

// Low-precision copy of a resident FP64 matrix, cast on the device.
template <class V>
static int cast_csr(DevCsr<V> &dst, const DevCsr<double> &src, hipStream_t s)
{
    dst.release();
    if (!src.present()) return CM_OK;
    dst.n = src.n;
    dst.nnz = src.nnz;
    dst.nblk = src.nblk;
    dst.tile = src.tile;
    dst.d16 = src.d16;
    dst.cbase = src.cbase;
    dst.slots = src.slots;  // slot-major and diagonal values are gathered when the format is chosen
    dst.dia = src.dia;
    // the staged 16-bit path pays off from a few nonzeros per row up (the
    // autotuner settles it per matrix by timing)
    dst.use16 = src.d16 && src.nnz >= 4LL * src.n;
    dst.owns_index = false;
    dst.ptr = src.ptr;
    dst.col = src.col;
    dst.blk = src.blk;
    dst.nlev = src.nlev;
    dst.lower = src.lower;
    dst.lrows = src.lrows;
    dst.loff = src.loff;
    dst.dloff = src.dloff;
    dst.lnnz = src.lnnz;
    dst.lpr = src.lpr;
    dst.tptr = src.tptr;
    dst.tcol = src.tcol;
    dst.tsrc = src.tsrc;
    dst.ntri = src.ntri;
    CM_TRY(dev_alloc(&dst.val, (size_t)dst.nnz + CSR_PAD));
    CM_HIP(hipMemsetAsync(dst.val + dst.nnz, 0, CSR_PAD * sizeof(V), s));
    if constexpr (std::is_same<V, __half>::value) {
        CM_TRY(dev_alloc(&dst.rowscale, (size_t)dst.n));
        k_half_rows<<<(dst.n + BS - 1) / BS, BS, 0, s>>>(dst.n, src.ptr, src.val, dst.val, dst.rowscale);
    } else {
        k_cast<V><<<long_grid(dst.nnz), BS, 0, s>>>(dst.nnz, src.val, dst.val);
    }
    if (dst.tsrc && dst.ntri) {  // the triangle's values in level order, from this copy's values
        CM_TRY(dev_alloc(&dst.tval, (size_t)dst.ntri));
        k_gather_vals<V><<<long_grid(dst.ntri), BS, 0, s>>>(dst.ntri, dst.tsrc, dst.val, dst.tval);
    }
    CM_HIP(hipGetLastError());
    return CM_OK;
}

// Block-relative 16-bit column indices of a resident FP64 matrix for the
// current row blocks (for the low-precision copies, which share them): each
// block's columns as offsets from its smallest column.  Left absent when a
// block's columns span 65536 or more.
static int build_idx16(DevCsr<double> &m, hipStream_t s)
{
    if (!m.present()) return CM_OK;
    (void)hipFree(m.d16);
    (void)hipFree(m.cbase);
    m.d16 = nullptr;
    m.cbase = nullptr;
    std::vector<int> ptr(m.n + 1), blk(m.nblk + 1), col(m.nnz);
    CM_HIP(hipMemcpyAsync(ptr.data(), m.ptr, ptr.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    CM_HIP(hipMemcpyAsync(blk.data(), m.blk, blk.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    if (m.nnz) CM_HIP(hipMemcpyAsync(col.data(), m.col, col.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    CM_HIP(hipStreamSynchronize(s));
    std::vector<int> cb(m.nblk, 0);
    std::vector<uint16_t> d((size_t)m.nnz + 2 * CSR_PAD, 0);
    for (int b = 0; b < m.nblk; b++) {
        const long long k0 = ptr[blk[b]], k1 = ptr[blk[b + 1]];
        int lo = INT32_MAX, hi = -1;
        for (long long k = k0; k < k1; k++) {
            lo = std::min(lo, col[k]);
            hi = std::max(hi, col[k]);
        }
        if (hi < 0) continue;
        if (hi - lo > 65535) return CM_OK;  // keep the 32-bit indices
        cb[b] = lo;
        for (long long k = k0; k < k1; k++) d[k] = (uint16_t)(col[k] - lo);
    }
    CM_TRY(dev_upload(&m.d16, d.data(), d.size(), s));
    CM_TRY(dev_upload(&m.cbase, cb.data(), cb.size(), s));
    CM_HIP(hipStreamSynchronize(s));
    return CM_OK;
}

// Slot-major index of a resident FP64 matrix (spmv_slot): blocks of
// SLOT_ROWS rows, rows ordered by length inside a block (longest first,
// ties in row order), nonzeros slot by slot.  Left absent when a row is
// longer than SLOT_MAXLEN.  The FP64 values are gathered into the new order.
static int build_slots(DevCsr<double> &m, hipStream_t s)
{
    if (!m.present() || m.slots) return CM_OK;
    const int n = m.n;
    std::vector<int> ptr(n + 1), col(m.nnz);
    CM_HIP(hipMemcpyAsync(ptr.data(), m.ptr, ptr.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    if (m.nnz) CM_HIP(hipMemcpyAsync(col.data(), m.col, col.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    CM_HIP(hipStreamSynchronize(s));
    SlotLayout L;
    if (!slot_layout(n, ptr.data(), col.data(), L)) return CM_OK;
    auto S = std::make_shared<SlotIndex>();
    S->n = n;
    S->nb = L.nb;
    S->nnz = m.nnz;
    S->nsoff = L.soff.size();
    CM_TRY(dev_upload(&S->hdr, L.hdr.data(), L.hdr.size(), s));
    CM_TRY(dev_upload(&S->soff, L.soff.data(), L.soff.size(), s));
    CM_TRY(dev_upload(&S->lrow, L.lrow.data(), L.lrow.size(), s));
    CM_TRY(dev_upload(&S->col, L.col.data(), L.col.size(), s));
    CM_TRY(dev_upload(&S->src, L.src.data(), L.src.size(), s));
    if (L.fits16) {
        CM_TRY(dev_upload(&S->d16, L.d16.data(), L.d16.size(), s));
        CM_TRY(dev_upload(&S->cbase, L.cbase.data(), L.cbase.size(), s));
    }
    CM_HIP(hipStreamSynchronize(s));  // host temporaries
    m.slots = S;
    return CM_OK;
}

// This copy's values in the slot-major order of its shared index.
template <class V>
static int build_slot_vals(DevCsr<V> &m, hipStream_t s)
{
    if (!m.present() || !m.slots || m.sval) return CM_OK;
    CM_TRY(dev_alloc(&m.sval, (size_t)m.nnz));
    if (m.nnz)
        k_gather_vals<V><<<long_grid(m.nnz), BS, 0, s>>>(m.nnz, m.slots->src, m.val, m.sval);
    CM_HIP(hipGetLastError());
    return CM_OK;
}

// ------------------------------------------- diagonal index on the device
//
// The layout dia_layout computes on the host, built from the resident
// matrix without copying it back: the candidate offsets are those of
// DIA_SAMPLE_ROWS sample rows spread by a multiplicative hash (row s * K mod
// n, K odd: distinct rows when n is a power of two, and no stride that lines
// up with a grid's boundary rows, which evenly spaced rows of a 2^k-row grid
// all are), so a diagonal held by 60% of the rows is missed only if its gaps
// fall on every sampled row; their
// entries are counted exactly (a histogram in LDS per workgroup), the
// selection follows dia_layout's rule, and the diagonal positions and the
// CSR remainder are placed by one thread per row in the entries' order.
// Matrices with a sampled row longer than DIA_SAMPLE_LEN, more than
// DIA_CAND_MAX candidate offsets, or remainders longer than SHORT_MAX (whose
// block products need the long-row tiles) are left to the host.
constexpr int DIA_SAMPLE_ROWS = 1024, DIA_SAMPLE_LEN = 64, DIA_CAND_MAX = 1024;

// offsets (column - row) of the S sample rows, DIA_SAMPLE_LEN slots per row
// (INT_MIN where the row ends); *over set when a sample row is longer
__global__ __launch_bounds__(BS) void k_dia_sample(int n, int S, const int *__restrict__ ptr,
                                                   const int *__restrict__ col, int *out, int *over)
{
    for (int q = blockIdx.x * BS + threadIdx.x; q < S * DIA_SAMPLE_LEN; q += gridDim.x * BS) {
        const int smp = q / DIA_SAMPLE_LEN, j = q % DIA_SAMPLE_LEN;
        const int i = S == n ? smp : (int)((unsigned long long)smp * 2654435761ULL % (unsigned long long)n);
        const int b = ptr[i], e = ptr[i + 1];
        out[q] = b + j < e ? col[b + j] - i : INT_MIN;
        if (j == 0 && e - b > DIA_SAMPLE_LEN) *over = 1;
    }
}

__global__ __launch_bounds__(BS) void k_fill_int(long long cnt, int v, int *a)
{
    for (long long q = (long long)blockIdx.x * BS + threadIdx.x; q < cnt; q += (long long)gridDim.x * BS) a[q] = v;
}

// map[list[q] + n - 1] = q
__global__ __launch_bounds__(BS) void k_dia_mark(int c, const int *__restrict__ list, int n, int *map)
{
    for (int q = blockIdx.x * BS + threadIdx.x; q < c; q += gridDim.x * BS) map[(size_t)(list[q] + n - 1)] = q;
}

// entries on each of the c marked offsets: LDS histogram per workgroup
__global__ __launch_bounds__(BS) void k_dia_hist(int n, const int *__restrict__ ptr, const int *__restrict__ col,
                                                 const int *__restrict__ map, int c, int *cnt)
{
    __shared__ int h[DIA_CAND_MAX];
    for (int q = threadIdx.x; q < c; q += BS) h[q] = 0;
    __syncthreads();
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS)
        for (int k = ptr[i]; k < ptr[i + 1]; k++) {
            const int q = map[(size_t)(col[k] - i + n - 1)];
            if (q >= 0) atomicAdd(&h[q], 1);
        }
    __syncthreads();
    for (int q = threadIdx.x; q < c; q += BS)
        if (h[q]) atomicAdd(&cnt[q], h[q]);
}

// dsrc[d * ldn + i] = k for the entries on diagonal d = map[offset] (the
// last one when a row repeats a column, as on the host); rcnt[i] = the row's
// other entries; *longrem set when a row keeps more than SHORT_MAX of them
__global__ __launch_bounds__(BS) void k_dia_place(int n, int ldn, const int *__restrict__ ptr,
                                                  const int *__restrict__ col, const int *__restrict__ map,
                                                  int *dsrc, int *rcnt, int *longrem)
{
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        int r = 0;
        for (int k = ptr[i]; k < ptr[i + 1]; k++) {
            const int d = map[(size_t)(col[k] - i + n - 1)];
            if (d >= 0)
                dsrc[(size_t)d * ldn + i] = k;
            else
                r++;
        }
        rcnt[i] = r;
        if (r > SHORT_MAX) *longrem = 1;
    }
}

// the remainder entries of every row, in order, from rptr[i]
__global__ __launch_bounds__(BS) void k_dia_rem(int n, const int *__restrict__ ptr, const int *__restrict__ col,
                                                const int *__restrict__ map, const int *__restrict__ rptr, int *rcol,
                                                int *rsrc)
{
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        int pos = rptr[i];
        for (int k = ptr[i]; k < ptr[i + 1]; k++)
            if (map[(size_t)(col[k] - i + n - 1)] < 0) {
                rcol[pos] = col[k];
                rsrc[pos] = k;
                pos++;
            }
    }
}

// Device temporaries freed together at scope exit.
struct DevTemps {
    std::vector<void *> p;
    ~DevTemps()
    {
        for (void *q : p) (void)hipFree(q);
    }
    template <class T>
    int alloc(T **d, size_t cnt)
    {
        CM_TRY(dev_alloc(d, cnt));
        p.push_back(*d);
        return CM_OK;
    }
};

// Device build of m's diagonal index; *done false when the matrix is left to
// the host build.  m.dia stays absent (done) when no diagonal qualifies.
static int build_dia_device(DevCsr<double> &m, hipStream_t s, bool *done)
{
    *done = false;
    const int n = m.n;
    if (n < 1 || m.nnz < 1 || 2LL * n - 1 > INT_MAX) return CM_OK;
    DevTemps tmp;
    const int S = std::min(n, DIA_SAMPLE_ROWS);
    int *smp, *flags;  // flags: [0] a long sample row, [1] a long remainder
    CM_TRY(tmp.alloc(&smp, (size_t)S * DIA_SAMPLE_LEN));
    CM_TRY(tmp.alloc(&flags, 2));
    CM_HIP(hipMemsetAsync(flags, 0, 2 * sizeof(int), s));
    k_dia_sample<<<(S * DIA_SAMPLE_LEN + BS - 1) / BS, BS, 0, s>>>(n, S, m.ptr, m.col, smp, flags);
    std::vector<int> hs((size_t)S * DIA_SAMPLE_LEN);
    int hf[2] = {0, 0};
    CM_HIP(hipMemcpyAsync(hs.data(), smp, hs.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    CM_HIP(hipMemcpyAsync(hf, flags, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    CM_HIP(hipStreamSynchronize(s));
    if (hf[0]) return CM_OK;
    std::vector<int> cand;
    for (int o : hs)
        if (o != INT_MIN) cand.push_back(o);
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    const int c = (int)cand.size();
    if (c == 0 || c > DIA_CAND_MAX) return CM_OK;
    // exact counts of the candidates
    const size_t nmap = 2 * (size_t)n - 1;
    int *map, *dl, *cnt;
    CM_TRY(tmp.alloc(&map, nmap));
    CM_TRY(tmp.alloc(&dl, (size_t)c));
    CM_TRY(tmp.alloc(&cnt, (size_t)c));
    k_fill_int<<<long_grid((long long)nmap), BS, 0, s>>>((long long)nmap, -1, map);
    CM_HIP(hipMemcpyAsync(dl, cand.data(), (size_t)c * sizeof(int), hipMemcpyHostToDevice, s));
    CM_HIP(hipMemsetAsync(cnt, 0, (size_t)c * sizeof(int), s));
    k_dia_mark<<<(c + BS - 1) / BS, BS, 0, s>>>(c, dl, n, map);
    k_dia_hist<<<std::min((n + BS - 1) / BS, 1024), BS, 0, s>>>(n, m.ptr, m.col, map, c, cnt);
    std::vector<int> hc((size_t)c);
    CM_HIP(hipMemcpyAsync(hc.data(), cnt, (size_t)c * sizeof(int), hipMemcpyDeviceToHost, s));
    CM_HIP(hipStreamSynchronize(s));
    // dia_layout's selection among the candidates
    std::vector<std::pair<long long, int>> counted;
    for (int q = 0; q < c; q++) counted.push_back({hc[q], cand[q]});
    const std::vector<int> off = dia_select(n, counted);
    *done = true;
    if (off.empty()) return CM_OK;
    const int nd = (int)off.size(), ldn = (n + DIA_ALIGN - 1) / DIA_ALIGN * DIA_ALIGN;
    auto D = std::make_shared<DiaIndex>();
    D->n = n;
    D->ldn = ldn;
    D->nd = nd;
    CM_TRY(dev_upload(&D->off, off.data(), off.size(), s));
    // positions on the selected diagonals (map: offset -> diagonal), remainder counts
    k_fill_int<<<long_grid((long long)nmap), BS, 0, s>>>((long long)nmap, -1, map);
    k_dia_mark<<<(nd + BS - 1) / BS, BS, 0, s>>>(nd, D->off, n, map);
    CM_TRY(dev_alloc(&D->dsrc, (size_t)nd * ldn));
    k_fill_int<<<long_grid((long long)nd * ldn), BS, 0, s>>>((long long)nd * ldn, -1, D->dsrc);
    int *rcnt, *rptr;
    CM_TRY(tmp.alloc(&rcnt, (size_t)n));
    CM_TRY(tmp.alloc(&rptr, (size_t)n + 1));
    CM_HIP(hipMemsetAsync(rptr, 0, sizeof(int), s));
    const int gr = std::min((n + BS - 1) / BS, 8192);
    k_dia_place<<<gr, BS, 0, s>>>(n, ldn, m.ptr, m.col, map, D->dsrc, rcnt, flags + 1);
    size_t tb = 0;
    CM_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tb, rcnt, rptr + 1, n, s));
    char *tbuf;
    CM_TRY(tmp.alloc(&tbuf, std::max<size_t>(tb, 1)));
    CM_HIP(hipcub::DeviceScan::InclusiveSum((void *)tbuf, tb, rcnt, rptr + 1, n, s));
    int tot = 0;
    CM_HIP(hipMemcpyAsync(&tot, rptr + n, sizeof(int), hipMemcpyDeviceToHost, s));
    CM_HIP(hipMemcpyAsync(hf, flags, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    CM_HIP(hipStreamSynchronize(s));
    if (hf[1]) {  // long remainders: the host build cuts them into tiles
        *done = false;
        return CM_OK;
    }
    D->nrem = tot;
    // no long rows: the tile fields as dia_layout leaves them
    D->nlong = 0;
    D->ntile = std::min(LONG_TILES_MAX, std::max(1, (n + LONG_TILE_COLS - 1) / LONG_TILE_COLS));
    D->tw = (n + D->ntile - 1) / D->ntile;
    D->nseg = 0;
    D->nlrem = 0;
    if (tot > 0) {
        D->rptr = rptr;
        tmp.p.erase(std::find(tmp.p.begin(), tmp.p.end(), (void *)rptr));  // kept by the index
        CM_TRY(dev_alloc(&D->rcol, (size_t)tot));
        CM_TRY(dev_alloc(&D->rsrc, (size_t)tot));
        k_dia_rem<<<gr, BS, 0, s>>>(n, m.ptr, m.col, map, D->rptr, D->rcol, D->rsrc);
    }
    CM_HIP(hipGetLastError());
    CM_HIP(hipStreamSynchronize(s));  // before the temporaries are freed
    m.dia = D;
    return CM_OK;
}

// Diagonal index of a resident FP64 matrix (spmv_dia): the diagonals held
// by at least DIA_FILL of the rows (the DIA_MAX most populated, in offset
// order), the other entries as a CSR remainder.  Left absent when no
// diagonal qualifies.  Built on the device where build_dia_device takes the
// matrix, else from a host copy (dia_layout).
static int build_dia(DevCsr<double> &m, hipStream_t s)
{
    if (!m.present() || m.dia) return CM_OK;
    bool done = false;
    CM_TRY(build_dia_device(m, s, &done));
    if (done) return CM_OK;
    const int n = m.n;
    std::vector<int> ptr(n + 1), col(m.nnz);
    CM_HIP(hipMemcpyAsync(ptr.data(), m.ptr, ptr.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    if (m.nnz) CM_HIP(hipMemcpyAsync(col.data(), m.col, col.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    CM_HIP(hipStreamSynchronize(s));
    DiaLayout L;
    if (!dia_layout(n, ptr.data(), col.data(), L)) return CM_OK;
    auto D = std::make_shared<DiaIndex>();
    D->n = n;
    D->ldn = L.ldn;
    D->nd = L.nd;
    D->nrem = (long long)L.rcol.size();
    D->nlong = L.nlong;
    D->ntile = L.ntile;
    D->tw = L.tw;
    D->nseg = L.nseg;
    D->nlrem = (long long)L.lcol.size();
    if (D->nlong) {
        CM_TRY(dev_upload(&D->lpos, L.lpos.data(), L.lpos.size(), s));
        CM_TRY(dev_upload(&D->lptr, L.lptr.data(), L.lptr.size(), s));
        CM_TRY(dev_upload(&D->lcol, L.lcol.data(), L.lcol.size(), s));
        CM_TRY(dev_upload(&D->lsrc, L.lsrc.data(), L.lsrc.size(), s));
    }
    CM_TRY(dev_upload(&D->off, L.off.data(), L.off.size(), s));
    CM_TRY(dev_upload(&D->dsrc, L.dsrc.data(), L.dsrc.size(), s));
    if (D->nrem) {
        CM_TRY(dev_upload(&D->rptr, L.rptr.data(), L.rptr.size(), s));
        CM_TRY(dev_upload(&D->rcol, L.rcol.data(), L.rcol.size(), s));
        CM_TRY(dev_upload(&D->rsrc, L.rsrc.data(), L.rsrc.size(), s));
    }
    CM_HIP(hipStreamSynchronize(s));  // host temporaries
    m.dia = D;
    return CM_OK;
}

// This copy's diagonal and remainder values (gathered from its CSR values).
template <class V>
static int build_dia_vals(DevCsr<V> &m, hipStream_t s)
{
    if (!m.present() || !m.dia || m.dval) return CM_OK;
    const DiaIndex &D = *m.dia;
    const long long nde = (long long)D.nd * D.ldn;
    CM_TRY(dev_alloc(&m.dval, (size_t)nde));
    CM_TRY(dev_alloc(&m.rval, (size_t)std::max<long long>(D.nrem, 1)));
    k_gather_vals<V><<<long_grid(nde), BS, 0, s>>>(nde, D.dsrc, m.val, m.dval);
    if (D.nrem) k_gather_vals<V><<<long_grid(D.nrem), BS, 0, s>>>(D.nrem, D.rsrc, m.val, m.rval);
    if (D.nlrem) {
        CM_TRY(dev_alloc(&m.lval, (size_t)D.nlrem));
        k_gather_vals<V><<<long_grid(D.nlrem), BS, 0, s>>>(D.nlrem, D.lsrc, m.val, m.lval);
    }
    CM_HIP(hipGetLastError());
    return CM_OK;
}

// Per level position q (row rows[q]): the count of its row's entries in the
// declared triangle (lower: j <= i, upper: j >= i).
__global__ __launch_bounds__(BS) void k_tri_count(int n, const int *__restrict__ rows, const int *__restrict__ ptr,
                                                  const int *__restrict__ col, bool lower, int *cnt)
{
    for (int q = blockIdx.x * BS + threadIdx.x; q < n; q += gridDim.x * BS) {
        const int i = rows[q];
        int c = 0;
        for (int k = ptr[i]; k < ptr[i + 1]; k++) c += lower ? col[k] <= i : col[k] >= i;
        cnt[q] = c;
    }
}

// The triangle's entries in level order: columns and CSR positions from tptr[q].
__global__ __launch_bounds__(BS) void k_tri_fill(int n, const int *__restrict__ rows, const int *__restrict__ ptr,
                                                 const int *__restrict__ col, bool lower,
                                                 const int *__restrict__ tptr, int *tcol, int *tsrc)
{
    for (int q = blockIdx.x * BS + threadIdx.x; q < n; q += gridDim.x * BS) {
        const int i = rows[q];
        int pos = tptr[q];
        for (int k = ptr[i]; k < ptr[i + 1]; k++) {
            const int j = col[k];
            if (lower ? j <= i : j >= i) {
                tcol[pos] = j;
                tsrc[pos] = k;
                pos++;
            }
        }
    }
}

// Dependency levels of a triangular factor (level_schedule, on the host):
// one launch of the triangular-solve kernel per level; the triangle's
// entries are then gathered in level order on the device.
static int build_levels(DevCsr<double> &m, const int *ptr, const int *col, bool lower, hipStream_t s)
{
    LevelSchedule L = level_schedule(m.n, ptr, col, lower, false);
    m.nlev = L.nlev;
    m.lower = lower;
    m.loff = L.loff;
    CM_TRY(dev_upload(&m.lrows, L.rows.data(), (size_t)m.n, s));
    CM_TRY(dev_upload(&m.dloff, L.loff.data(), L.loff.size(), s));
    // the triangle's entries in level order
    const int n = m.n;
    DevTemps tmp;
    int *cnt;
    CM_TRY(tmp.alloc(&cnt, (size_t)n));
    CM_TRY(dev_alloc(&m.tptr, (size_t)n + 1));
    CM_HIP(hipMemsetAsync(m.tptr, 0, sizeof(int), s));
    const int g = std::min((n + BS - 1) / BS, 8192);
    k_tri_count<<<g, BS, 0, s>>>(n, m.lrows, m.ptr, m.col, lower, cnt);
    size_t tb = 0;
    CM_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tb, cnt, m.tptr + 1, n, s));
    char *tbuf;
    CM_TRY(tmp.alloc(&tbuf, std::max<size_t>(tb, 1)));
    CM_HIP(hipcub::DeviceScan::InclusiveSum((void *)tbuf, tb, cnt, m.tptr + 1, n, s));
    std::vector<int> tp((size_t)n + 1);
    CM_HIP(hipMemcpyAsync(tp.data(), m.tptr, tp.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    CM_HIP(hipStreamSynchronize(s));
    m.ntri = tp[n];
    m.lnnz.assign(L.nlev, 0);  // the triangle's entries each level's rows hold (read by its launch)
    for (int l = 0; l < L.nlev; l++) m.lnnz[l] = tp[L.loff[l + 1]] - tp[L.loff[l]];
    CM_TRY(dev_alloc(&m.tcol, (size_t)m.ntri));
    CM_TRY(dev_alloc(&m.tsrc, (size_t)m.ntri));
    CM_TRY(dev_alloc(&m.tval, (size_t)m.ntri));
    k_tri_fill<<<g, BS, 0, s>>>(n, m.lrows, m.ptr, m.col, lower, m.tptr, m.tcol, m.tsrc);
    if (m.ntri) k_gather_vals<double><<<long_grid(m.ntri), BS, 0, s>>>(m.ntri, m.tsrc, m.val, m.tval);
    // lanes per row: the power of two at or above the triangle's mean row length, 1..16
    const double mean = m.n ? (double)m.ntri / m.n : 0.0;
    m.lpr = 1;
    while (m.lpr < 16 && m.lpr < mean) m.lpr *= 2;
    CM_HIP(hipGetLastError());
    CM_HIP(hipStreamSynchronize(s));  // before the temporaries are freed
    return CM_OK;
}

// maxmsteps per package (AFIIM/src/bicgstab.cpp:114, PDFIIM/src/bicgstab.cpp:106,
// ACIFIM/src/bicgstab.cpp:132).
static int max_div_steps(int variant, int n)
{
    if (variant == 2) return std::min(std::max(n, 2000) / 50, 40);
    return std::min(std::max(n, 500) / 50, 20);
}

// Passes per captured hipGraph (besides the one-pass graph).
constexpr int GRAPH_PASSES = 4;

// Kernel categories of the per-kernel timing (cm_system_kernel_stats).
enum KernelCat { KC_A = 0, KC_F1 = 1, KC_F2 = 2, KC_LU = 3, KC_VEC = 4, KC_SETUP = 5, KC_FIN = 6, KC_N = 7 };

// Optional per-launch timing: an event pair around every launch of a chosen
// category, tagged with the BiCGSTAB pass it belongs to and its algorithmic
// bytes.  Launches the device skipped (after the exit) are dropped when the
// solve ends, so the averages cover executed kernels only.
struct Prof {
    bool on = false;
    unsigned mask = 0;  // categories whose launches are bracketed
    int sample = 0;     // > 0: only the first `sample` passes of a solve are bracketed (and launched
                        // one by one); the others replay from the captured graphs
    int sampled = 0;    // passes of the current solve launched with brackets so far
    int pass = -1;      // pass being enqueued (-1: setup)
    bool p_update = false;
    struct Rec {
        int cat, pass;
        bool p_update;
        double bytes, fbytes;
        hipEvent_t e0, e1;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    long long launches[KC_N] = {};
    double ms[KC_N] = {}, bytes[KC_N] = {}, fbytes[KC_N] = {};

    hipEvent_t get()
    {
        if (pool.empty()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            return e;
        }
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    void reset_stats()
    {
        for (int c = 0; c < KC_N; c++) {
            launches[c] = 0;
            ms[c] = 0.0;
            bytes[c] = 0.0;
            fbytes[c] = 0.0;
        }
    }
    // Fold the finished solve's records into the totals.  A pass executes
    // fully iff it computed its residual (pass < ntrace); its p update runs
    // only when the loop went on (pass < ntrace - 1).
    void collect(int ntrace)
    {
        for (Rec &r : recs) {
            bool ran = (r.pass < 0) || (r.pass < ntrace && (!r.p_update || r.pass < ntrace - 1));
            if (ran && r.e0 && r.e1) {
                float t = 0.f;
                if (hipEventElapsedTime(&t, r.e0, r.e1) == hipSuccess) {
                    launches[r.cat]++;
                    ms[r.cat] += t;
                    bytes[r.cat] += r.bytes;
                    fbytes[r.cat] += r.fbytes;
                }
            }
            if (r.e0) pool.push_back(r.e0);
            if (r.e1) pool.push_back(r.e1);
        }
        recs.clear();
    }
    ~Prof()
    {
        for (Rec &r : recs) {
            if (r.e0) (void)hipEventDestroy(r.e0);
            if (r.e1) (void)hipEventDestroy(r.e1);
        }
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
};

// Host-side description of a system (CSR arrays owned by the caller).
struct HostSys {
    int variant, n, a_symmetric, nr;
    char mode;
    const int *Ai, *Aj;
    const double *Av;
    const int *Gi, *Gj;
    const double *Gv, *IDv;
    const int *Hi, *Hj;
    const double *Hv;
    const int *Li, *Lj;
    const double *Lv;
    const int *Ui, *Uj;
    const double *Uv;
};

// The BiCGSTAB core for vectors of type T, A values of type VA and factor
// values of type VF.
template <class T, class VA, class VF>
struct Krylov {
    int variant = 0, n = 0, nr = 1;
    char mode = 'E';
    bool sym = false;
    double eps = 1.11022302462516e-16;  // stagnation constant, AFIIM/src/bicgstab.cpp:108
    hipStream_t stream = nullptr;       // owned by the System
    DevCsr<VA> A;
    DevCsr<VF> G, H, GT, L, U;
    T *idv = nullptr;
    T *b = nullptr, *x = nullptr, *r0 = nullptr, *r = nullptr, *p = nullptr, *v = nullptr, *s = nullptr,
      *t = nullptr, *pc = nullptr, *sc = nullptr, *tm = nullptr, *t1 = nullptr, *t2 = nullptr, *t3 = nullptr;
    SolveState *st = nullptr;   // device
    SolveState *hst = nullptr;  // pinned host, 3 slots (init + 2 polling slots)
    double *partials = nullptr;
    size_t partials_cap = 0;  // doubles
    double *trace = nullptr;
    int trace_cap_dev = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
    Prof prof;
    // passes replayed from captured hipGraphs (1 pass, GRAPH_PASSES passes)
    // when no launch is being bracketed by profiling events
    bool use_graph = true;
    hipGraphExec_t g1 = nullptr, gN = nullptr;

    void drop_graphs()
    {
        if (g1) (void)hipGraphExecDestroy(g1);
        if (gN) (void)hipGraphExecDestroy(gN);
        g1 = gN = nullptr;
    }

    ~Krylov()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graphs();
        T *bufs[] = {idv, b, x, r0, r, p, v, s, t, pc, sc, tm, t1, t2, t3};
        for (T *q : bufs) (void)hipFree(q);
        (void)hipFree(partials);
        (void)hipFree(trace);
        (void)hipFree(st);
        if (hst) (void)hipHostFree(hst);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
    }

    // device bytes held: matrices, 15 n-vectors, reduction partials, trace
    size_t bytes() const
    {
        return A.bytes() + G.bytes() + H.bytes() + GT.bytes() + L.bytes() + U.bytes() +
               15 * (size_t)n * sizeof(T) + (size_t)std::max(A.nblk, MAX_RED_BLOCKS) * 4 * sizeof(double) +
               (size_t)trace_cap_dev * sizeof(double);
    }

    bool uses_gh() const { return variant != 2 || mode == 'E' || mode == 'C'; }
    bool uses_lu() const { return variant == 2 && ((mode == 'C' && nr > 1) || mode == 'I' || mode == 'J'); }
    // L and U in sparse products (mode C's sweeps), not only in triangular solves
    bool lu_spmv() const { return variant == 2 && mode == 'C' && nr > 1; }

    // vectors, state, partials, events
    int alloc_work(int nblk_a)
    {
        T **vecs[] = {&b, &x, &r0, &r, &p, &v, &s, &t, &pc, &sc, &tm, &t1, &t2, &t3};
        for (T **q : vecs) CM_TRY(dev_alloc(q, n));
        CM_HIP(hipMemsetAsync(x, 0, (size_t)n * sizeof(T), stream));
        CM_HIP(hipMemsetAsync(b, 0, (size_t)n * sizeof(T), stream));
        CM_HIP(hipMalloc((void **)&st, sizeof(SolveState)));
        CM_HIP(hipHostMalloc((void **)&hst, 3 * sizeof(SolveState), hipHostMallocDefault));
        const int maxblk = std::max(nblk_a, MAX_RED_BLOCKS);
        CM_HIP(hipMalloc((void **)&partials, (size_t)maxblk * 4 * sizeof(double)));
        partials_cap = (size_t)maxblk * 4;
        for (auto &e : ev) CM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        return CM_OK;
    }

    // the reduction partials must hold K <= 4 values per SpMV workgroup
    int ensure_partials(int nblk)
    {
        const size_t need = (size_t)std::max(nblk, MAX_RED_BLOCKS) * 4;
        if (need <= partials_cap) return CM_OK;
        CM_HIP(hipStreamSynchronize(stream));
        drop_graphs();  // their launches carry the old buffer
        (void)hipFree(partials);
        partials = nullptr;
        CM_HIP(hipMalloc((void **)&partials, need * sizeof(double)));
        partials_cap = need;
        return CM_OK;
    }

    void copy_shape(int var, int nn, int nr_, char mode_, bool sym_, hipStream_t strm)
    {
        variant = var;
        n = nn;
        nr = nr_;
        mode = mode_;
        sym = sym_;
        stream = strm;
    }

    // Launch f on the stream, bracketed by timing events when profiling.
    // bytes: the launch's algorithmic traffic; fbytes: what the storage
    // format in use touches once (the same unless a format reads more or
    // less than plain CSR; < 0 means the same).
    template <class F>
    void timed(int cat, double bytes, F f, double fbytes = -1.0)
    {
        if (!prof.on || !(prof.mask & (1u << cat))) {
            f();
            return;
        }
        Prof::Rec rec{cat, prof.pass, prof.p_update, bytes, fbytes < 0.0 ? bytes : fbytes, prof.get(), prof.get()};
        if (rec.e0) (void)hipEventRecord(rec.e0, stream);
        f();
        if (rec.e1) (void)hipEventRecord(rec.e1, stream);
        prof.recs.push_back(rec);
    }

    // Algorithmic bytes of one SpMV launch, independent of the storage
    // format: the CSR matrix once (values, column indices, row pointers, and
    // the per-row scales of FP16 values), x once, y written once, plus the
    // epilogue's extra n-vectors.
    template <class V, class Epi>
    double alg_bytes(const DevCsr<V> &m) const
    {
        return (double)m.nnz * (sizeof(V) + 4.0) + 4.0 * (m.n + 1) + (m.rowscale ? 4.0 * m.n : 0.0) +
               (double)sizeof(T) * m.n * (2 + Epi::XV);
    }

    // Compulsory bytes of one SpMV launch in the format in use: its matrix
    // arrays once (row blocks, slot tables or diagonals with their zeros
    // included), x once, y written once, plus the epilogue's extra n-vectors.
    template <class V, class Epi>
    double spmv_bytes(const DevCsr<V> &m) const
    {
        if (m.use_dia && m.dval) {  // diagonals (zeros included), offsets, remainder once
            const DiaIndex &D = *m.dia;
            return (double)D.nd * m.n * sizeof(V) + 4.0 * D.nd + (double)D.nrem * (sizeof(V) + 4.0) +
                   (D.rptr ? 4.0 * (m.n + 1) : 0.0) + (m.rowscale ? 4.0 * m.n : 0.0) +
                   (double)sizeof(T) * m.n * (2 + Epi::XV);
        }
        if (m.use_slot && m.sval) {  // slot-major: block headers, slot tables and row order once
            const SlotIndex &S = *m.slots;
            return (double)m.nnz * (sizeof(V) + (m.slot16 ? 2.0 : 4.0)) +
                   4.0 * (5.0 * S.nb + (double)S.nsoff + S.n) + (m.slot16 ? 4.0 * S.nb : 0.0) +
                   (m.rowscale ? 4.0 * m.n : 0.0) + (double)sizeof(T) * m.n * (2 + Epi::XV);
        }
        const bool i16 = m.d16 && m.use16 && !std::is_same<V, double>::value;
        return (double)m.nnz * (sizeof(V) + (i16 ? 2.0 : 4.0)) + 4.0 * (m.n + 1) + 8.0 * (m.nblk + 1) +
               (i16 ? 4.0 * m.nblk : 0.0) + (m.rowscale ? 4.0 * m.n : 0.0) + (double)sizeof(T) * m.n * (2 + Epi::XV);
    }

    template <class V, class Epi>
    void spmv(const DevCsr<V> &m, const T *xin, Epi epi, bool use_state, int cat, SolveState *fst = nullptr)
    {
        if (m.nblk == 0) return;
        using Acc = typename std::conditional<std::is_same<T, double>::value, double, float>::type;
        // fst: a separate state (refinement residual) instead of the solve's
        SolveState *kst = fst ? fst : st;
        SolveState *gate = use_state ? kst : nullptr;
        if (m.use_dia && m.dval) {
            const DiaIndex &D = *m.dia;
            constexpr int RPT = dia_rpt<V>();
            const int nb = m.dia2 ? (m.n + BS * 2 * RPT - 1) / (BS * 2 * RPT) : (m.n + BS * RPT - 1) / (BS * RPT);
            const bool rem_lds = (double)D.nrem >= DIA_REM_LDS_MIN * (double)m.n;
            timed(cat, alg_bytes<V, Epi>(m), [&] {
                auto go = [&](auto k) {
                    k<<<nb, BS, 0, stream>>>(m.n, D.ldn, D.nd, D.off, m.dval, D.rptr, D.rcol, m.rval, m.rowscale, xin,
                                             epi, gate, partials);
                };
                if (m.dia2) {
                    if (rem_lds) go(spmv_dia<V, T, Acc, Epi, 9, 2 * RPT, true>);
                    else go(spmv_dia<V, T, Acc, Epi, 9, 2 * RPT, false>);
                } else {
                    if (rem_lds) go(spmv_dia<V, T, Acc, Epi, 8, RPT, true>);
                    else go(spmv_dia<V, T, Acc, Epi, 8, RPT, false>);
                }
            }, spmv_bytes<V, Epi>(m));
            if constexpr (Epi::K > 0)
                timed(KC_FIN, 8.0 * Epi::K * nb,
                      [&] { k_finalize<Epi::K, Epi><<<1, FBS, 0, stream>>>(nb, partials, kst, epi); });
            return;
        }
        if (m.use_slot && m.sval) {
            const SlotIndex &S = *m.slots;
            timed(cat, alg_bytes<V, Epi>(m), [&] {
                if (m.slot16)
                    spmv_slot<V, T, Acc, Epi, true><<<S.nb, BS, 0, stream>>>(
                        S.hdr, S.soff, S.lrow, S.col, S.d16, S.cbase, m.sval, m.rowscale, xin, epi, gate, partials);
                else
                    spmv_slot<V, T, Acc, Epi, false><<<S.nb, BS, 0, stream>>>(
                        S.hdr, S.soff, S.lrow, S.col, S.d16, S.cbase, m.sval, m.rowscale, xin, epi, gate, partials);
            }, spmv_bytes<V, Epi>(m));
            if constexpr (Epi::K > 0)
                timed(KC_FIN, 8.0 * Epi::K * S.nb,
                      [&] { k_finalize<Epi::K, Epi><<<1, FBS, 0, stream>>>(S.nb, partials, kst, epi); });
            return;
        }
        timed(cat, alg_bytes<V, Epi>(m), [&] {
            if constexpr (!std::is_same<V, double>::value) {
                if (m.d16 && m.use16) {  // low precision with 16-bit indices
                    if (m.tile == 1024)
                        spmv_vl16<1024, V, T, Acc, Epi><<<m.nblk, BS, 0, stream>>>(
                            m.ptr, m.d16, m.cbase, m.val, m.rowscale, m.blk, xin, epi, gate, partials);
                    else if (m.tile == 4096)
                        spmv_vl16<4096, V, T, Acc, Epi><<<m.nblk, BS, 0, stream>>>(
                            m.ptr, m.d16, m.cbase, m.val, m.rowscale, m.blk, xin, epi, gate, partials);
                    else
                        spmv_vl16<2048, V, T, Acc, Epi><<<m.nblk, BS, 0, stream>>>(
                            m.ptr, m.d16, m.cbase, m.val, m.rowscale, m.blk, xin, epi, gate, partials);
                    return;
                }
            }
            if constexpr (std::is_same<V, __half>::value) {
                // 2-byte values: the vector-load variant (fewer load instructions)
                if (m.tile == 1024)
                    spmv_vl<1024, V, T, Acc, Epi><<<m.nblk, BS, 0, stream>>>(m.ptr, m.col, m.val, m.rowscale,
                                                                            m.blk, xin, epi, gate, partials);
                else if (m.tile == 4096)
                    spmv_vl<4096, V, T, Acc, Epi><<<m.nblk, BS, 0, stream>>>(m.ptr, m.col, m.val, m.rowscale,
                                                                            m.blk, xin, epi, gate, partials);
                else
                    spmv_vl<2048, V, T, Acc, Epi><<<m.nblk, BS, 0, stream>>>(m.ptr, m.col, m.val, m.rowscale,
                                                                            m.blk, xin, epi, gate, partials);
                return;
            }
            if (m.tile == 1024)
                spmv_kernel<1024, V, T, Acc, Epi><<<m.nblk, BS, 0, stream>>>(m.ptr, m.col, m.val, m.rowscale,
                                                                            m.blk, xin, epi, gate, partials);
            else if (m.tile == 4096)
                spmv_kernel<4096, V, T, Acc, Epi><<<m.nblk, BS, 0, stream>>>(m.ptr, m.col, m.val, m.rowscale,
                                                                            m.blk, xin, epi, gate, partials);
            else
                spmv_kernel<2048, V, T, Acc, Epi><<<m.nblk, BS, 0, stream>>>(m.ptr, m.col, m.val, m.rowscale,
                                                                            m.blk, xin, epi, gate, partials);
        }, spmv_bytes<V, Epi>(m));
        if constexpr (Epi::K > 0)
            timed(KC_FIN, 8.0 * Epi::K * m.nblk,
                  [&] { k_finalize<Epi::K, Epi><<<1, FBS, 0, stream>>>(m.nblk, partials, kst, epi); });
    }

    // x = T^{-1} b for a unit triangular factor with its level schedule: one
    // launch per level, or one workgroup for a run of at least two consecutive
    // levels that each take at most TRSV_RUN_STEPS steps of its lane groups
    void trsv(const DevCsr<VF> &m, const T *bin, T *xout, bool use_state)
    {
        const SolveState *gate = use_state ? st : nullptr;
        const int run_rows = TRSV_RUN_STEPS * (TRSV_RUN_THREADS / m.lpr);
        auto small = [&](int l) { return m.loff[l + 1] - m.loff[l] <= run_rows; };
        auto lanes = [&](auto go) {
            switch (m.lpr) {
            case 1: go(std::integral_constant<int, 1>{}); break;
            case 2: go(std::integral_constant<int, 2>{}); break;
            case 4: go(std::integral_constant<int, 4>{}); break;
            case 8: go(std::integral_constant<int, 8>{}); break;
            default: go(std::integral_constant<int, 16>{}); break;
            }
        };
        // the levels' entries (value + column), row list and row pointers, and per row the
        // right-hand side read, the solution written and (once over all levels) read back
        auto level_bytes = [&](int l) {
            const int cnt = m.loff[l + 1] - m.loff[l];
            return (double)m.lnnz[l] * (sizeof(VF) + 4.0) + (double)cnt * (4.0 + 8.0 + 3.0 * sizeof(T));
        };
        for (int l = 0; l < m.nlev;) {
            int l2 = l;
            while (l2 < m.nlev && small(l2)) l2++;
            if (l2 - l >= 2 && m.dloff) {
                double by = 0.0;
                for (int q = l; q < l2; q++) by += level_bytes(q);
                timed(KC_LU, by, [&] {
                    lanes([&](auto lz) {
                        constexpr int LPR = decltype(lz)::value;
                        k_trsv_run<T, VF, LPR><<<1, TRSV_RUN_THREADS, 0, stream>>>(
                            l, l2, m.dloff, m.lrows, m.tptr, m.tcol, m.tval, m.rowscale, bin, xout, gate);
                    });
                });
                l = l2;
                continue;
            }
            const int cnt = m.loff[l + 1] - m.loff[l];
            const int q0 = m.loff[l];
            timed(KC_LU, level_bytes(l), [&] {
                lanes([&](auto lz) {
                    constexpr int LPR = decltype(lz)::value;
                    const int g = (int)(((long long)cnt * LPR + BS - 1) / BS);
                    k_trsv_level<T, VF, LPR><<<g, BS, 0, stream>>>(cnt, m.lrows + q0, m.tptr + q0, m.tcol, m.tval,
                                                                   m.rowscale, bin, xout, gate);
                });
            });
            l++;
        }
    }

    // out = M in, per the reference's apply (ACIFIM/src/bicgstab.cpp:372-428);
    // for AFIIM / PDFIIM: out = G (IDv .* (H in)), H = G^T when symmetric.
    void enqueue_apply(const T *in, T *out, bool use_state)
    {
        if (variant != 2 || mode == 'E') {
            const DevCsr<VF> &Hm = sym ? GT : H;
            spmv(Hm, in, EpiScale<T, T>{tm, idv}, use_state, KC_F1);
            spmv(G, tm, EpiStore<T>{out}, use_state, KC_F2);
            return;
        }
        if (mode == 'I' || mode == 'J') {
            // I: t1 = L^{-1} b; t2 = t1 .* IDv; x = U^{-1} t2.   J: x = U^{-1} L^{-1} b.
            trsv(L, in, t1, use_state);
            if (mode == 'I') {
                const int g = vec_grid(n);
                timed(KC_VEC, 3.0 * sizeof(T) * n, [&] {
                    k_vmul<T><<<g, BS, 0, stream>>>(n, t1, idv, t2, use_state ? st : nullptr);
                });
                trsv(U, t2, out, use_state);
            } else {
                trsv(U, t1, out, use_state);
            }
            return;
        }
        // mode C: t1 = H b; repeat: t1 += H (b - L t1); t2 = t1 .* IDv;
        //         x = G t2; repeat: x += G (t2 - U x)
        if (nr <= 1) {
            spmv(H, in, EpiScale<T, T>{t2, idv}, use_state, KC_F1);
        } else {
            spmv(H, in, EpiStore<T>{t1}, use_state, KC_F1);
            for (int it = 1; it < nr; it++) {
                spmv(L, t1, EpiSubFrom<T>{t3, in}, use_state, KC_LU);
                if (it + 1 < nr) spmv(H, t3, EpiAccum<T>{t1}, use_state, KC_F1);
                else spmv(H, t3, EpiAccumScale<T, T>{t1, t2, idv}, use_state, KC_F1);
            }
        }
        spmv(G, t2, EpiStore<T>{out}, use_state, KC_F2);
        for (int it = 1; it < nr; it++) {
            spmv(U, out, EpiSubFrom<T>{t3, t2}, use_state, KC_LU);
            spmv(G, t3, EpiAccum<T>{out}, use_state, KC_F2);
        }
    }

    void enqueue_iteration()
    {
        const int g = vec_grid(n);
        const double nn = (double)n, w = (double)sizeof(T);
        prof.pass++;
        enqueue_apply(p, pc, true);
        spmv(A, pc, EpiAlpha<T>{v, r0}, true, KC_A);
        timed(KC_VEC, 3.0 * w * nn, [&] { k_s_update<T><<<g, BS, 0, stream>>>(n, r, v, s, st); });
        enqueue_apply(s, sc, true);
        spmv(A, sc, EpiOmega<T>{t, s}, true, KC_A);
        timed(KC_VEC, 9.0 * w * nn, [&] {
            k_xr_update<T><<<g, BS, 0, stream>>>(n, x, pc, sc, r, s, t, r0, st, partials);
        });
        timed(KC_FIN, 32.0 * g,
              [&] { k_finalize<4, FinPass><<<1, FBS, 0, stream>>>(g, partials, st, FinPass{}); });
        prof.p_update = true;
        timed(KC_VEC, 4.0 * w * nn, [&] { k_p_update<T><<<g, BS, 0, stream>>>(n, p, r, v, st); });
        prof.p_update = false;
    }

    int capture(hipGraphExec_t &out, int passes)
    {
        hipGraph_t graph = nullptr;
        const bool prof_on = prof.on;  // captured passes carry no timing events
        const int pass = prof.pass;
        prof.on = false;
        CM_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < passes; i++) enqueue_iteration();
        const hipError_t cerr = hipStreamEndCapture(stream, &graph);
        prof.on = prof_on;
        prof.pass = pass;
        CM_HIP(cerr);
        const hipError_t err = hipGraphInstantiate(&out, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        CM_HIP(err);
        return CM_OK;
    }

    int enqueue_passes(int size)
    {
        if (!use_graph || (prof.on && prof.sample <= 0)) {
            for (int i = 0; i < size; i++) enqueue_iteration();
            return CM_OK;
        }
        if (prof.on && prof.sampled < prof.sample) {  // the solve's first passes, bracketed
            const int e = std::min(size, prof.sample - prof.sampled);
            for (int i = 0; i < e; i++) enqueue_iteration();
            prof.sampled += e;
            size -= e;
            if (size == 0) return CM_OK;
        }
        if (!g1) {
            CM_TRY(capture(g1, 1));
            CM_TRY(capture(gN, GRAPH_PASSES));
        }
        for (; size >= GRAPH_PASSES; size -= GRAPH_PASSES) CM_HIP(hipGraphLaunch(gN, stream));
        for (; size > 0; size--) CM_HIP(hipGraphLaunch(g1, stream));
        return CM_OK;
    }

    // Writes the initial state (done = 0) of a solve of NMAX passes.
    int begin(double tol, int NMAX, int info_in, int trace_cap)
    {
        if (trace_cap > 0 && trace_cap > trace_cap_dev) {
            (void)hipFree(trace);
            trace = nullptr;
            CM_HIP(hipMalloc((void **)&trace, (size_t)trace_cap * sizeof(double)));
            trace_cap_dev = trace_cap;
        }
        SolveState init;
        std::memset(&init, 0, sizeof init);
        init.tol = tol;
        init.eps = eps;
        init.nmax = NMAX;
        init.maxmsteps = max_div_steps(variant, n);
        init.variant = variant;
        init.info = info_in;
        init.trace = trace_cap > 0 ? trace : nullptr;
        init.trace_cap = trace_cap > 0 ? trace_cap : 0;
        hst[0] = init;
        CM_HIP(hipMemcpyAsync(st, &hst[0], sizeof(SolveState), hipMemcpyHostToDevice, stream));
        prof.pass = -1;
        prof.sampled = 0;
        return CM_OK;
    }

    // r = b - A x, r0 = p = r (x as given)
    void setup_from_x() { spmv(A, x, EpiSetup<T>{r, r0, p, b}, true, KC_SETUP); }

    // x = 0, r = r0 = p = b
    void setup_zero()
    {
        const int g = vec_grid(n);
        timed(KC_SETUP, 5.0 * sizeof(T) * n,
              [&] { k_setup_zero<T><<<g, BS, 0, stream>>>(n, b, x, r, r0, p, st, partials); });
        timed(KC_FIN, 8.0 * g,
              [&] { k_finalize<1, FinSetup><<<1, FBS, 0, stream>>>(g, partials, st, FinSetup{}); });
    }

    // Runs passes until the device state says done.  Passes are enqueued in
    // chunks, at most two chunks outstanding, and the state is read back at
    // the end of each chunk.  The next chunk's size is the number of passes
    // the residual's recent contraction rate says are still needed (at least
    // 2, at most 32), so few no-op passes are enqueued past the exit.
    int run(int NMAX, SolveState &fin_state)
    {
        int qs[2];
        int nq = 0, slot = 0, last_chunk = 4;
        long long enq = 0;
        auto push = [&](int size) -> int {
            CM_TRY(enqueue_passes(size));
            CM_HIP(hipMemcpyAsync(&hst[1 + slot], st, sizeof(SolveState), hipMemcpyDeviceToHost, stream));
            CM_HIP(hipEventRecord(ev[slot], stream));
            enq += size;
            qs[nq++] = slot;
            slot ^= 1;
            last_chunk = size;
            return CM_OK;
        };
        const long long nmax_passes = std::max(NMAX, 0);
        CM_TRY(push((int)std::min<long long>(4, std::max<long long>(nmax_passes, 1))));
        bool have_prev = false;
        int prev_iter = -1;      // the snapshot before the latest one
        double prev_rc = 0.0;
        for (;;) {
            if (nq < 2 && enq < nmax_passes) {
                // Passes still needed, predicted from the faster of the mean
                // contraction rate since the start and the rate over the last
                // chunk (a residual that plateaus and then drops fast would
                // otherwise keep the chunks long past the exit); half of them
                // are enqueued now, so the chunks shrink towards the exit.
                // Without a contraction yet, the chunk doubles up to 8.
                long long want = std::min(2 * last_chunk, 8);
                if (have_prev) {
                    const SolveState &sn = fin_state;  // latest snapshot
                    double rate = 2.0;
                    if (sn.iter >= 2 && sn.rc > 0.0 && sn.ro > 0.0 && sn.rc < sn.ro)
                        rate = std::pow(sn.rc / sn.ro, 1.0 / sn.iter);
                    if (prev_iter >= 0 && sn.iter > prev_iter && sn.rc > 0.0 && prev_rc > sn.rc)
                        rate = std::min(rate, std::pow(sn.rc / prev_rc, 1.0 / (sn.iter - prev_iter)));
                    if (rate < 1.0 && rate > 0.0) {
                        const double need = std::log(sn.tol * sn.ro / sn.rc) / std::log(rate);
                        const long long ahead = enq - sn.iter;
                        want = (long long)std::max(1.0, std::min(16.0, std::ceil(0.5 * (need - (double)ahead))));
                    }
                }
                want = std::min(want, nmax_passes - enq);
                if (want > 0) CM_TRY(push((int)want));
            }
            if (nq == 0) return fail(CM_ERR_HIP, "solve did not reach an exit");
            const int o = qs[0];
            qs[0] = qs[1];
            nq--;
            CM_HIP(hipEventSynchronize(ev[o]));
            const SolveState snap = hst[1 + o];
            if (have_prev) {
                prev_iter = fin_state.iter;
                prev_rc = fin_state.rc;
            }
            fin_state = snap;
            have_prev = true;
            if (snap.done) break;
        }
        CM_HIP(hipStreamSynchronize(stream));
        CM_HIP(hipGetLastError());
        if (prof.on) prof.collect(fin_state.ntrace);
        if (!fin_state.done) return fail(CM_ERR_HIP, "solve state did not reach done");
        return CM_OK;
    }
};

using K64 = Krylov<double, double, double>;
using K32 = Krylov<float, float, float>;
using K16 = Krylov<float, float, __half>;

static int create_k64(K64 &k, const HostSys &h, hipStream_t strm)
{
    const char mode = (h.variant == 2) ? h.mode : 'E';
    k.copy_shape(h.variant, h.n, h.nr, mode, h.variant == 1 && h.a_symmetric, strm);
    const int n = h.n;
    // column ranges of the matrices whose columns the host does not read
    // again: checked on the device, one flag read back at the end
    struct Flag {
        int *d = nullptr;
        ~Flag() { (void)hipFree(d); }
    } bad;
    CM_TRY(dev_alloc(&bad.d, 1));
    CM_HIP(hipMemsetAsync(bad.d, 0, sizeof(int), strm));
    if (k.sym) {  // the expansion and the transpose read the columns on the host
        CM_TRY(check_csr(n, h.Ai, h.Aj));
        CM_TRY(check_csr(n, h.Gi, h.Gj));
        std::vector<int> fp, fc;
        std::vector<double> fv;
        host_sym_expand(n, h.Ai, h.Aj, h.Av, fp, fc, fv);
        CM_TRY(upload_csr(k.A, n, fp.data(), fc.data(), fv.data(), strm));
        std::vector<int> tp, tc;
        std::vector<double> tv;
        host_transpose(n, h.Gi, h.Gj, h.Gv, tp, tc, tv);
        CM_TRY(upload_csr(k.GT, n, tp.data(), tc.data(), tv.data(), strm));
    } else {
        CM_TRY(upload_csr(k.A, n, h.Ai, h.Aj, h.Av, strm, bad.d));
        if (k.uses_gh()) CM_TRY(upload_csr(k.H, n, h.Hi, h.Hj, h.Hv, strm, bad.d));
    }
    if (k.uses_gh()) CM_TRY(upload_csr(k.G, n, h.Gi, h.Gj, h.Gv, strm, k.sym ? nullptr : bad.d));
    if (k.uses_lu()) {
        const bool levels = k.mode == 'I' || k.mode == 'J';  // the level schedule reads the columns on the host
        CM_TRY(upload_csr(k.L, n, h.Li, h.Lj, h.Lv, strm, levels ? nullptr : bad.d));
        CM_TRY(upload_csr(k.U, n, h.Ui, h.Uj, h.Uv, strm, levels ? nullptr : bad.d));
        if (levels) {
            CM_TRY(build_levels(k.L, h.Li, h.Lj, true, strm));
            CM_TRY(build_levels(k.U, h.Ui, h.Uj, false, strm));
        }
    }
    CM_TRY(dev_upload(&k.idv, h.IDv, (size_t)n, strm));
    CM_TRY(k.alloc_work(k.A.nblk));
    int hbad = 0;
    CM_HIP(hipMemcpyAsync(&hbad, bad.d, sizeof(int), hipMemcpyDeviceToHost, strm));
    CM_HIP(hipStreamSynchronize(strm));
    if (hbad) return fail(CM_ERR_ARG, CSR_COL_RANGE);
    return CM_OK;
}

// Low-precision inner solver cast from the resident FP64 one.
template <class KI>
static int create_inner(KI &ki, K64 &k)
{
    // 16-bit indices for the low-precision products (built once, shared)
    for (DevCsr<double> *m : {&k.A, &k.G, &k.H, &k.GT, &k.L, &k.U})
        if (m->present() && !m->d16) CM_TRY(build_idx16(*m, k.stream));
    ki.copy_shape(k.variant, k.n, k.nr, k.mode, k.sym, k.stream);
    ki.eps = 5.9604644775390625e-08;  // 2^-24: the stagnation test at FP32 resolution
    CM_TRY(cast_csr(ki.A, k.A, k.stream));
    CM_TRY(cast_csr(ki.G, k.G, k.stream));
    CM_TRY(cast_csr(ki.H, k.H, k.stream));
    CM_TRY(cast_csr(ki.GT, k.GT, k.stream));
    CM_TRY(cast_csr(ki.L, k.L, k.stream));
    CM_TRY(cast_csr(ki.U, k.U, k.stream));
    CM_TRY(dev_alloc(&ki.idv, (size_t)k.n));
    k_cast<float><<<long_grid(k.n), BS, 0, k.stream>>>(k.n, k.idv, ki.idv);
    CM_TRY(ki.alloc_work(ki.A.nblk));
    CM_HIP(hipStreamSynchronize(k.stream));
    return CM_OK;
}


// Inner tolerance of one refinement step: inner_tol, or looser when the
// residual rn only has to shrink by less than that to reach tol * res0 (half
// of the remaining factor, at most 1e-2).  The last step of a refinement then
// stops well before a full inner solve.  A loose step that fails to reduce
// the FP64 residual (the low-precision correction is limited by its own
// rounding there) is followed by a step at inner_tol instead of ending the
// refinement.
static inline double adaptive_inner_tol(double inner_tol, double tol, double res0, double rn)
{
    if (!(rn > 0.0) || !(res0 > 0.0)) return inner_tol;
    const double need = 0.5 * tol * res0 / rn;
    return std::max(inner_tol, std::min(need, 1e-2));
}

// ------------------------------------------------------------- autotuner
//
// Row blocks of a resident FP64 matrix re-cut for `tile` staged nonzeros per
// SpMV workgroup.  The row pointers are read back from the device.
static int retile(DevCsr<double> &m, int tile, hipStream_t s)
{
    if (!m.present() || m.tile == tile) return CM_OK;
    std::vector<int> ptr(m.n + 1);
    CM_HIP(hipMemcpyAsync(ptr.data(), m.ptr, (size_t)(m.n + 1) * sizeof(int), hipMemcpyDeviceToHost, s));
    CM_HIP(hipStreamSynchronize(s));
    int nblk = 0;
    std::vector<int> blk = block_table(m.n, ptr.data(), &nblk, tile);
    int *nb = nullptr;
    CM_TRY(dev_upload(&nb, blk.data(), blk.size(), s));
    CM_HIP(hipStreamSynchronize(s));
    (void)hipFree(m.blk);
    m.blk = nb;
    m.nblk = nblk;
    m.tile = tile;
    return CM_OK;
}

// A low-precision copy shares the FP64 matrix's index arrays: follow its
// row blocks and its slot-major index.
template <class V>
static void follow_blocks(DevCsr<V> &dst, const DevCsr<double> &src)
{
    if (!dst.present()) return;
    dst.blk = src.blk;
    dst.nblk = src.nblk;
    dst.tile = src.tile;
    dst.d16 = src.d16;
    dst.cbase = src.cbase;
    if (!dst.d16) dst.use16 = false;
    if (dst.slots != src.slots) {
        (void)hipFree(dst.sval);
        dst.sval = nullptr;
        dst.slots = src.slots;
        dst.use_slot = dst.slot16 = false;
    }
    if (dst.dia != src.dia) {
        (void)hipFree(dst.dval);
        (void)hipFree(dst.rval);
        (void)hipFree(dst.lval);
        dst.dval = dst.rval = dst.lval = nullptr;
        dst.dia = src.dia;
        dst.use_dia = dst.block_dia = false;
    }
}

// Selects the SpMV format of one resident copy (SpmvFormat).  The staged
// 16-bit form exists for the low-precision copies when the 16-bit offsets
// were built; the slot-major forms when the shared slot-major index was
// (FP64 matrices whose rows are at most SLOT_MAXLEN long), the 16-bit one
// when its blocks' column spans allow.  Slot-major values are gathered on
// first use.
template <class V>
static int set_format(DevCsr<V> &m, int f, hipStream_t s)
{
    if (!m.present()) return fail(CM_ERR_ARG, "matrix not resident");
    const bool low = !std::is_same<V, double>::value;
    switch (f) {
    case FMT_STAGED:
        m.use_slot = m.use_dia = false;
        m.use16 = false;
        return CM_OK;
    case FMT_STAGED16:
        if (!low || !m.d16) return fail(CM_ERR_ARG, "staged 16-bit-index format not available for this copy");
        m.use_slot = m.use_dia = false;
        m.use16 = true;
        return CM_OK;
    case FMT_SLOT:
    case FMT_SLOT16:
        if (!m.slots) return fail(CM_ERR_ARG, "slot-major format not available (rows too long or index not built)");
        if (f == FMT_SLOT16 && !m.slots->d16)
            return fail(CM_ERR_ARG, "slot-major 16-bit offsets not available (column span)");
        CM_TRY(build_slot_vals(m, s));
        m.use_dia = false;
        m.use_slot = true;
        m.slot16 = f == FMT_SLOT16;
        return CM_OK;
    case FMT_DIA:
    case FMT_DIA2:
        if (!m.dia) return fail(CM_ERR_ARG, "diagonal format not available (no diagonal held by enough rows)");
        CM_TRY(build_dia_vals(m, s));
        m.use_slot = false;
        m.use_dia = true;
        m.dia2 = f == FMT_DIA2;
        return CM_OK;
    default:
        return fail(CM_ERR_ARG, "unknown SpMV format");
    }
}

template <class V>
static int get_format(const DevCsr<V> &m)
{
    if (!m.present()) return -1;
    if (m.use_dia && m.dval) return m.dia2 ? FMT_DIA2 : FMT_DIA;
    if (m.use_slot && m.sval) return m.slot16 ? FMT_SLOT16 : FMT_SLOT;
    return (!std::is_same<V, double>::value && m.d16 && m.use16) ? FMT_STAGED16 : FMT_STAGED;
}

// Average time of reps products of m on the Krylov's stream (one warm-up).
template <class KI, class V>
static int time_product(KI &ki, const DevCsr<V> &m, int reps, hipEvent_t e0, hipEvent_t e1, float *ms)
{
    using T = typename std::remove_pointer<decltype(ki.x)>::type;
    ki.spmv(m, ki.x, EpiStore<T>{ki.t}, false, KC_A);
    CM_HIP(hipEventRecord(e0, ki.stream));
    for (int i = 0; i < reps; i++) ki.spmv(m, ki.x, EpiStore<T>{ki.t}, false, KC_A);
    CM_HIP(hipEventRecord(e1, ki.stream));
    CM_HIP(hipEventSynchronize(e1));
    CM_HIP(hipEventElapsedTime(ms, e0, e1));
    return CM_OK;
}

// Autotuner, per SpMV matrix of one Krylov (A, G, H, GT; L and U in mode
// C's sweeps): every available format timed (reps products each), the
// fastest kept.  The staged row
// blocks are the FP64 matrix's current ones.
template <class KI>
static int tune_formats(KI &ki, int reps, hipEvent_t e0, hipEvent_t e1)
{
    auto one = [&](auto &m) -> int {
        if (!m.present()) return CM_OK;
        if (m.slots) CM_TRY(ki.ensure_partials(m.slots->nb));
        if (m.dia) CM_TRY(ki.ensure_partials((m.n + BS - 1) / BS));  // at least the launch's workgroups
        float best = 1e30f;
        int pick = FMT_STAGED;
        for (int f = FMT_STAGED; f <= FMT_DIA2; f++) {
            if (f == FMT_DIA2 && m.dia && m.dia->nd < DIA2_MIN_ND) continue;
            if (set_format(m, f, ki.stream) != CM_OK) continue;
            float ms = 0.f;
            CM_TRY(time_product(ki, m, reps, e0, e1, &ms));
            if (ms < best) {
                best = ms;
                pick = f;
            }
        }
        return set_format(m, pick, ki.stream);
    };
    CM_TRY(one(ki.A));
    CM_TRY(one(ki.G));
    CM_TRY(one(ki.H));
    CM_TRY(one(ki.GT));
    if (ki.lu_spmv()) {
        CM_TRY(one(ki.L));
        CM_TRY(one(ki.U));
    }
    ki.drop_graphs();
    return CM_OK;
}

// ------------------------------------------------------- block (multi-RHS)

// Row bins of one matrix for the SpMM kernels: rows with at most SHORT_MAX
// nonzeros (KB lanes each) and longer rows (a wave each).
struct RowBins {
    int *shortr = nullptr, *longr = nullptr;
    int ns = 0, nl = 0;
    int short_max = SHORT_MAX;  // rows longer than this go to the wave-per-row kernel
    ~RowBins()
    {
        (void)hipFree(shortr);
        (void)hipFree(longr);
    }
};

template <class V>
static int make_bins(RowBins &b, const DevCsr<V> &m, hipStream_t s, int short_max = SHORT_MAX)
{
    if (!m.present()) return CM_OK;
    std::vector<int> ptr(m.n + 1);
    CM_HIP(hipMemcpyAsync(ptr.data(), m.ptr, (size_t)(m.n + 1) * sizeof(int), hipMemcpyDeviceToHost, s));
    CM_HIP(hipStreamSynchronize(s));
    std::vector<int> sh, lo;
    row_bins(m.n, ptr.data(), short_max, sh, lo);
    (void)hipFree(b.shortr);
    (void)hipFree(b.longr);
    b.shortr = b.longr = nullptr;
    b.short_max = short_max;
    b.ns = (int)sh.size();
    b.nl = (int)lo.size();
    // every row short: the list is 0..n-1 and the kernel takes q for rows[q]
    if (b.nl > 0) CM_TRY(dev_upload(&b.shortr, sh.data(), sh.size(), s));
    CM_TRY(dev_upload(&b.longr, lo.data(), lo.size(), s));
    CM_HIP(hipStreamSynchronize(s));
    return CM_OK;
}

// Block BiCGSTAB over KB interleaved columns, on matrices owned by a Krylov
// of the same precision (A, G and H or G^T; apply mode E).
//
// Iterations are enqueued in chunks (the per-column states say when every
// column is done).  With graphs on, one pass and GRAPH_PASSES passes are
// captured once into hipGraphs (kernel nodes only) and a chunk is a few graph
// launches: the host cost of a pass drops from ~16 kernel launches to a
// fraction of one graph launch, which is what small systems are bound by.
template <class T, class VA, class VF, int KB>
struct BlockKrylov {
    int n = 0, variant = 0, ncols = 0;
    double eps = 1.11022302462516e-16;
    hipStream_t stream = nullptr;
    const DevCsr<VA> *A = nullptr;
    const DevCsr<VF> *G = nullptr, *Hm = nullptr;
    const T *idv = nullptr;
    RowBins bA, bG, bH;
    T *b = nullptr, *x = nullptr, *r0 = nullptr, *r = nullptr, *p = nullptr, *v = nullptr, *s = nullptr,
      *t = nullptr, *pc = nullptr, *sc = nullptr, *tm = nullptr;
    SolveState *st = nullptr, *hst = nullptr;  // KB device states; pinned host copies (3 * KB)
    BlockFlags *fl = nullptr, *hfl = nullptr;  // device flags; pinned host copies (3)
    double *partials = nullptr, *partials2 = nullptr;  // slot partials; first-stage sums
    size_t partial_slots = 0;
    T *racc = nullptr;  // long-row remainder sums of the diagonal block products (KB per long row)
    size_t racc_cap = 0;
    bool racc_short = false;  // a product was refused: racc smaller than it needs (nothing launched)
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool use_graph = true;
    hipGraphExec_t g1 = nullptr, gN = nullptr;  // one pass, GRAPH_PASSES passes

    ~BlockKrylov()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        if (g1) (void)hipGraphExecDestroy(g1);
        if (gN) (void)hipGraphExecDestroy(gN);
        T *bufs[] = {b, x, r0, r, p, v, s, t, pc, sc, tm, racc};
        for (T *q : bufs) (void)hipFree(q);
        (void)hipFree(st);
        (void)hipFree(fl);
        (void)hipFree(partials);
        (void)hipFree(partials2);
        if (hst) (void)hipHostFree(hst);
        if (hfl) (void)hipHostFree(hfl);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
    }

    size_t ne() const { return (size_t)n * KB; }
    int vgrid() const { return (int)std::max<size_t>(1, std::min<size_t>((ne() + BS - 1) / BS, MAX_RED_BLOCKS)); }

    // per-column totals of nb partial slots -> fin on the states fst
    template <int K, class Fin>
    void finalize(int nb, Fin fin, SolveState *fst, bool gate, int update_all)
    {
        const double *src = partials;
        if (nb > FIN_ONE) {
            const int g1 = (nb + FIN_PER - 1) / FIN_PER;
            k_block_sum1<K, KB><<<g1, FBS, 0, stream>>>(nb, FIN_PER, partials, partials2, gate ? fl : nullptr);
            src = partials2;
            nb = g1;
        }
        k_block_finalize<K, KB, Fin><<<1, FBS, 0, stream>>>(nb, src, fst, fl, fin, update_all);
    }
    // SpMM workgroups walk contiguous chunks of the row lists; the grid is
    // one step per workgroup up to SPMM_CAP workgroups.  Column reductions
    // over more than FIN_ONE partial slots go through a first stage of
    // FIN_PER slots per workgroup (k_block_sum1), so the one-workgroup
    // finalize always reads a few hundred slots at most.
    static constexpr int SROWS = BS / ColSplit<T, KB>::LPRW;  // short rows per workgroup step
    static constexpr int SPMM_CAP = 1 << 16;
    static constexpr int FIN_ONE = 256, FIN_PER = 64;
    static int short_grid(const RowBins &bn) { return std::min((bn.ns + SROWS - 1) / SROWS, SPMM_CAP); }
    static int long_grid(const RowBins &bn) { return std::min((bn.nl + (BS / WAVE) - 1) / (BS / WAVE), SPMM_CAP); }
    static int dia_grid(int nrows) { return std::min((nrows + SROWS - 1) / SROWS, SPMM_CAP); }

    // Elements of the long-row remainder sums one product of m needs: the
    // row sums, then the segment sums when its remainders are tiled.
    template <class V>
    static size_t racc_of(const DevCsr<V> *m)
    {
        if (!m || !m->block_dia || !m->dia || !m->dia->nlong) return 0;
        const DiaIndex &D = *m->dia;
        return (size_t)D.nlong * KB * (D.ntile > 1 ? 1 + (size_t)D.ntile : 1);
    }
    size_t racc_need() const { return std::max(racc_of(A), std::max(racc_of(G), racc_of(Hm))); }

    // Room for the long-row remainder sums of the matrices that use the
    // diagonal block products.  Called outside graph capture before every
    // solve, timing and tuning pass (a matrix may have switched to the
    // diagonal block products since the workspace was made).
    int ensure_racc()
    {
        const size_t need = racc_need();
        if (need <= racc_cap) return CM_OK;
        CM_HIP(hipStreamSynchronize(stream));
        drop_graphs();  // their launches carry the old buffer
        (void)hipFree(racc);
        racc = nullptr;
        CM_HIP(hipMalloc((void **)&racc, need * sizeof(T)));
        racc_cap = need;
        return CM_OK;
    }

    // device bytes of a block solver of this type for an n-row system whose
    // three matrices have (short + long) row-bin workgroups `slots` and whose
    // long-row remainder sums take `racc` elements
    static size_t bytes_for(int n, size_t slots, size_t racc)
    {
        const size_t e = (size_t)n * KB;
        return 11 * e * sizeof(T) + KB * sizeof(SolveState) + sizeof(BlockFlags) +
               std::max(slots, (size_t)MAX_RED_BLOCKS) * 4 * KB * sizeof(double) + 3 * (size_t)n * sizeof(int) +
               racc * sizeof(T);
    }
    size_t bytes() const { return bytes_for(n, partial_slots, racc_cap); }

    template <class KR>
    int create(const KR &k, hipStream_t strm)
    {
        n = k.n;
        variant = k.variant;
        stream = strm;
        eps = k.eps;
        A = &k.A;
        G = &k.G;
        Hm = k.sym ? &k.GT : &k.H;
        idv = k.idv;
        CM_TRY(make_bins(bA, *A, stream));
        CM_TRY(make_bins(bG, *G, stream));
        CM_TRY(make_bins(bH, *Hm, stream));
        T **vecs[] = {&b, &x, &r0, &r, &p, &v, &s, &t, &pc, &sc, &tm};
        for (T **q : vecs) CM_TRY(dev_alloc(q, ne()));
        CM_HIP(hipMemsetAsync(x, 0, ne() * sizeof(T), stream));
        CM_HIP(hipMemsetAsync(b, 0, ne() * sizeof(T), stream));
        CM_HIP(hipMalloc((void **)&st, KB * sizeof(SolveState)));
        CM_HIP(hipHostMalloc((void **)&hst, 3 * KB * sizeof(SolveState), hipHostMallocDefault));
        CM_HIP(hipMalloc((void **)&fl, sizeof(BlockFlags)));
        CM_HIP(hipHostMalloc((void **)&hfl, 3 * sizeof(BlockFlags), hipHostMallocDefault));
        size_t slots = std::max((size_t)vgrid(), (size_t)dia_grid(n));
        for (const RowBins *bn : {&bA, &bG, &bH})
            slots = std::max(slots, (size_t)(short_grid(*bn) + long_grid(*bn)));
        partial_slots = std::max(slots, (size_t)MAX_RED_BLOCKS);
        CM_HIP(hipMalloc((void **)&partials, partial_slots * 4 * KB * sizeof(double)));
        CM_HIP(hipMalloc((void **)&partials2, (partial_slots / FIN_PER + 1) * 4 * KB * sizeof(double)));
        for (auto &e : ev) CM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        CM_TRY(ensure_racc());
        CM_HIP(hipStreamSynchronize(stream));
        return CM_OK;
    }

    // y = M x (+ epilogue and, for reducing epilogues, the per-column finalize
    // on the states fst)
    template <class V, class Epi>
    void spmm(const DevCsr<V> &m, const RowBins &bn, const T *xin, Epi epi, SolveState *fst, bool gate = true)
    {
        using Acc = typename std::conditional<std::is_same<T, double>::value, double, float>::type;
        const BlockFlags *gf = gate ? fl : nullptr;
        if (m.block_dia && m.dval) {
            const DiaIndex &D = *m.dia;
            const int gd = dia_grid(n);
            if (racc_of(&m) > racc_cap) {  // never launch through a missing or short buffer
                racc_short = true;
                return;
            }
            if (D.nlong) {  // long remainders first: segment sums, added by tile into racc
                const long long nseg = D.nseg;
                const int gseg = (int)std::min<long long>((nseg + SROWS - 1) / SROWS, SPMM_CAP);
                T *seg = D.ntile > 1 ? racc + (size_t)D.nlong * KB : racc;
                spmm_short<V, T, Acc, KB, BEpiStore<T>><<<gseg, BS, 0, stream>>>(
                    (int)nseg, nullptr, D.lptr, D.lcol, m.lval, nullptr, xin, BEpiStore<T>{seg}, gf, partials, 0);
                if (D.ntile > 1) {
                    const long long ne_l = (long long)D.nlong * KB;
                    const int gt = (int)std::min<long long>((ne_l + BS - 1) / BS, 8192);
                    k_long_tiles<T, Acc><<<gt, BS, 0, stream>>>(ne_l, D.ntile, seg, racc, gf);
                }
            }
            auto go = [&](auto k) {
                k<<<gd, BS, 0, stream>>>(n, D.ldn, D.nd, D.off, m.dval, D.rptr, D.rcol, m.rval,
                                         D.nlong ? D.lpos : nullptr, racc, m.rowscale, xin, epi, gf, partials, 0);
            };
            // the remainder the kernel itself walks: long rows' entries arrive summed in racc
            if ((double)(D.nrem - D.nlrem) >= DIA_REM_LDS_MIN * (double)n) go(spmm_dia<V, T, Acc, KB, Epi, true>);
            else go(spmm_dia<V, T, Acc, KB, Epi, false>);
            if constexpr (Epi::K > 0) finalize<Epi::K>(gd, epi, fst, gate, gate ? 1 : 0);
            return;
        }
        const int gs = short_grid(bn), gl = long_grid(bn);
        if (gs > 0)
            spmm_short<V, T, Acc, KB, Epi><<<gs, BS, 0, stream>>>(bn.ns, bn.shortr, m.ptr, m.col, m.val, m.rowscale,
                                                                 xin, epi, gf, partials, 0);
        if (gl > 0)
            spmm_long<V, T, Acc, KB, Epi><<<gl, BS, 0, stream>>>(bn.nl, bn.longr, m.ptr, m.col, m.val, m.rowscale,
                                                                xin, epi, gf, partials, gs);
        if constexpr (Epi::K > 0) finalize<Epi::K>(gs + gl, epi, fst, gate, gate ? 1 : 0);
    }

    // Average time (microseconds) of `reps` back-to-back products y = M x of
    // matrix 'A', 'G' or 'H' (H: the first factor, G^T on the symmetric
    // branch) over the KB columns, bracketed by events on this stream; bytes =
    // compulsory traffic of one product (matrix and row lists once, x and y
    // once per column).
    // error when a product was refused for want of racc room (clears the flag)
    int check_racc()
    {
        if (!racc_short) return CM_OK;
        racc_short = false;
        drop_graphs();
        return fail(CM_ERR_HIP, "block product refused: long-row buffer smaller than the matrix needs");
    }

    template <class V>
    int time_one(const DevCsr<V> &m, const RowBins &bn, int reps, double *avg_us, double *bytes)
    {
        CM_TRY(ensure_racc());
        hipEvent_t e0, e1;
        CM_HIP(hipEventCreate(&e0));
        CM_HIP(hipEventCreate(&e1));
        spmm(m, bn, p, BEpiStore<T>{v}, nullptr, false);
        CM_HIP(hipEventRecord(e0, stream));
        for (int i = 0; i < reps; i++) spmm(m, bn, p, BEpiStore<T>{v}, nullptr, false);
        CM_HIP(hipEventRecord(e1, stream));
        CM_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        CM_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        CM_TRY(check_racc());
        *avg_us = 1e3 * ms / reps;
        if (m.block_dia && m.dval) {  // diagonals (zeros included) and remainder once, X and Y once
            const DiaIndex &D = *m.dia;
            // (+ lpos, the segment pointers, and the segment sums
            // written and read back when the long remainders are tiled)
            *bytes = (double)D.nd * m.n * sizeof(V) + 4.0 * D.nd + (double)D.nrem * (sizeof(V) + sizeof(int)) +
                     (D.rptr ? 4.0 * (m.n + 1) : 0.0) + (D.nlong ? 4.0 * m.n + 4.0 * D.nseg : 0.0) +
                     (D.ntile > 1 ? 2.0 * D.nseg * KB * sizeof(T) : 0.0) +
                     2.0 * (double)ne() * sizeof(T) + (m.rowscale ? 4.0 * m.n : 0.0);
        } else {
            *bytes = (double)m.nnz * (sizeof(V) + sizeof(int)) + 4.0 * (m.n + 1) +
                     4.0 * ((bn.shortr ? bn.ns : 0) + bn.nl) +
                     2.0 * (double)ne() * sizeof(T) + (m.rowscale ? 4.0 * m.n : 0.0);
        }
        CM_HIP(hipGetLastError());
        return CM_OK;
    }

    // Autotuner of the row binning: for each matrix, the short-row limit
    // (16, 32, 64 or 128 nonzeros) whose SpMM over the KB columns is fastest
    // (reps products per candidate, events on this stream).  Captured graphs
    // are dropped (their launches carry the old bins).
    template <class V>
    int tune_one(const DevCsr<V> &m, RowBins &bn, int reps, int *chosen)
    {
        *chosen = 0;
        if (!m.present()) return CM_OK;
        double best = 1e300;
        int best_t = SHORT_MAX;
        for (int thr : {16, 32, 64, 128}) {
            CM_TRY(make_bins(bn, m, stream, thr));
            CM_TRY(fit_partials());
            double us = 0.0, by = 0.0;
            CM_TRY(time_one(m, bn, reps, &us, &by));
            if (us < best) {
                best = us;
                best_t = thr;
            }
        }
        CM_TRY(make_bins(bn, m, stream, best_t));
        *chosen = best_t;
        return CM_OK;
    }

    int fit_partials()
    {
        size_t slots = std::max((size_t)vgrid(), (size_t)dia_grid(n));
        for (const RowBins *bn : {&bA, &bG, &bH})
            slots = std::max(slots, (size_t)(short_grid(*bn) + long_grid(*bn)));
        if (slots <= partial_slots) return CM_OK;
        CM_HIP(hipStreamSynchronize(stream));
        (void)hipFree(partials);
        (void)hipFree(partials2);
        partials = partials2 = nullptr;
        CM_HIP(hipMalloc((void **)&partials, slots * 4 * KB * sizeof(double)));
        CM_HIP(hipMalloc((void **)&partials2, (slots / FIN_PER + 1) * 4 * KB * sizeof(double)));
        partial_slots = slots;
        return CM_OK;
    }

    void drop_graphs()
    {
        if (g1) (void)hipGraphExecDestroy(g1);
        if (gN) (void)hipGraphExecDestroy(gN);
        g1 = gN = nullptr;
    }

    int tune(int reps, int *chosen /* A, G, H */)
    {
        drop_graphs();
        CM_TRY(tune_one(*A, bA, reps, &chosen[0]));
        CM_TRY(tune_one(*G, bG, reps, &chosen[1]));
        CM_TRY(tune_one(*Hm, bH, reps, &chosen[2]));
        return CM_OK;
    }

    int time_spmm(char which, int reps, double *avg_us, double *bytes)
    {
        if (which == 'A') return time_one(*A, bA, reps, avg_us, bytes);
        if (which == 'G') return time_one(*G, bG, reps, avg_us, bytes);
        if (which == 'H') return time_one(*Hm, bH, reps, avg_us, bytes);
        return fail(CM_ERR_ARG, "matrix must be 'A', 'G' or 'H'");
    }

    void enqueue_apply(const T *in, T *out)
    {
        spmm(*Hm, bH, in, BEpiScale<T, T, KB>{tm, idv}, st);
        spmm(*G, bG, tm, BEpiStore<T>{out}, st);
    }

    void enqueue_iteration()
    {
        const int g = vgrid();
        enqueue_apply(p, pc);
        spmm(*A, bA, pc, BEpiAlpha<T>{v, r0}, st);
        k_block_s<T, KB><<<g, BS, 0, stream>>>(ne(), r, v, s, st, fl);
        enqueue_apply(s, sc);
        spmm(*A, bA, sc, BEpiOmega<T>{t, s}, st);
        k_block_xr<T, KB><<<g, BS, 0, stream>>>(ne(), x, pc, sc, r, s, t, r0, st, fl, partials);
        finalize<4>(g, FinPass{}, st, true, 1);
        k_block_p<T, KB><<<g, BS, 0, stream>>>(ne(), p, r, v, st, fl);
    }

    int capture(hipGraphExec_t &out, int passes)
    {
        hipGraph_t graph = nullptr;
        CM_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < passes; i++) enqueue_iteration();
        CM_HIP(hipStreamEndCapture(stream, &graph));
        const hipError_t err = hipGraphInstantiate(&out, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        CM_HIP(err);
        return CM_OK;
    }

    int enqueue_passes(int size)
    {
        if (!use_graph) {
            for (int i = 0; i < size; i++) enqueue_iteration();
            CM_TRY(check_racc());
            CM_HIP(hipGetLastError());
            return CM_OK;
        }
        if (!g1) {
            CM_TRY(capture(g1, 1));
            CM_TRY(capture(gN, GRAPH_PASSES));
            CM_TRY(check_racc());
        }
        for (; size >= GRAPH_PASSES; size -= GRAPH_PASSES) CM_HIP(hipGraphLaunch(gN, stream));
        for (; size > 0; size--) CM_HIP(hipGraphLaunch(g1, stream));
        return CM_OK;
    }

    // initial per-column states; columns not in `active` start done
    int begin(double tol, int NMAX, const int *active, const double *col_tol = nullptr)
    {
        CM_TRY(ensure_racc());
        for (int j = 0; j < KB; j++) {
            SolveState &q = hst[j];
            std::memset(&q, 0, sizeof q);
            q.tol = (col_tol && j < ncols) ? col_tol[j] : tol;
            q.eps = eps;
            q.nmax = NMAX;
            q.maxmsteps = max_div_steps(variant, n);
            q.variant = variant;
            q.done = (j < ncols && (!active || active[j])) ? 0 : 1;
        }
        hfl[0].all_done = 0;
        hfl[0].ncols = ncols;
        CM_HIP(hipMemcpyAsync(st, hst, KB * sizeof(SolveState), hipMemcpyHostToDevice, stream));
        CM_HIP(hipMemcpyAsync(fl, hfl, sizeof(BlockFlags), hipMemcpyHostToDevice, stream));
        return CM_OK;
    }

    void setup_from_x() { spmm(*A, bA, x, BEpiSetup<T>{r, r0, p, b}, st); }

    void setup_zero()
    {
        const int g = vgrid();
        k_block_setup_zero<T, KB><<<g, BS, 0, stream>>>(ne(), b, x, r, r0, p, partials);
        finalize<1>(g, FinSetup{}, st, true, 1);
    }

    // passes until every column is done; fin receives the final states.
    // Chunks of passes are enqueued ahead (at most two outstanding, the next
    // one sized from the residual contraction of the slowest live column).
    int run(int NMAX, SolveState *fin)
    {
        CM_TRY(check_racc());  // the setup product
        int qs[2], nq = 0, slot = 0, chunk = 4;
        long long enq = 0;
        const long long nmax_passes = std::max(NMAX, 0);
        auto push = [&](int size) -> int {
            CM_TRY(enqueue_passes(size));
            CM_HIP(hipMemcpyAsync(&hst[KB * (1 + slot)], st, KB * sizeof(SolveState), hipMemcpyDeviceToHost, stream));
            CM_HIP(hipMemcpyAsync(&hfl[1 + slot], fl, sizeof(BlockFlags), hipMemcpyDeviceToHost, stream));
            CM_HIP(hipEventRecord(ev[slot], stream));
            enq += size;
            qs[nq++] = slot;
            slot ^= 1;
            return CM_OK;
        };
        CM_TRY(push((int)std::min<long long>(4, std::max<long long>(nmax_passes, 1))));
        for (;;) {
            if (nq < 2 && enq < nmax_passes) CM_TRY(push((int)std::min<long long>(chunk, nmax_passes - enq)));
            if (nq == 0) return fail(CM_ERR_HIP, "block solve did not reach an exit");
            const int o = qs[0];
            qs[0] = qs[1];
            nq--;
            CM_HIP(hipEventSynchronize(ev[o]));
            if (hfl[1 + o].all_done) {
                std::memcpy(fin, &hst[KB * (1 + o)], KB * sizeof(SolveState));
                break;
            }
            // passes still needed by the slowest live column at its current
            // contraction rate, assuming the same rate from here on
            double need = 4.0;
            for (int j = 0; j < ncols; j++) {
                const SolveState &q = hst[KB * (1 + o) + j];
                if (q.done || q.iter < 2 || !(q.rc > 0.0) || !(q.ro > 0.0)) continue;
                const double rate = std::pow(q.rc / q.ro, 1.0 / q.iter);
                if (rate < 1.0 && rate > 0.0) need = std::max(need, std::log(q.tol * q.ro / q.rc) / std::log(rate));
            }
            const long long ahead = enq - hst[KB * (1 + o)].iter;  // passes already queued beyond this snapshot
            chunk = (int)std::min<double>(16.0, std::max<double>(1.0, std::ceil(0.5 * (need - (double)std::max<long long>(ahead, 0)))));
        }
        CM_HIP(hipStreamSynchronize(stream));
        CM_HIP(hipGetLastError());
        return CM_OK;
    }
};

// Type-erased block solver of one block width.
struct BlockBase {
    int kb = 0, ncols = 0;
    const K32 *k32 = nullptr;  // inner-solve matrices, once mixed precision is enabled
    const K16 *k16 = nullptr;
    virtual ~BlockBase() {}
    virtual int set(int nc, const double *B, const double *X0) = 0;
    virtual int get(double *X) = 0;
    virtual int solve64(double tol, int NMAX, double *resval, int *iter, int *info) = 0;
    virtual int solve_mixed(int bits, double tol, double inner_tol, int NMAX, int max_outer, double *resval,
                            int *outer, int *inner_total, int *info) = 0;
    virtual void set_stream(hipStream_t s) = 0;
    virtual void set_graphs(bool on) = 0;
    virtual void drop_graphs() = 0;
    virtual int ensure_workspace() = 0;
    virtual size_t bytes() const = 0;
    virtual int time_spmm(char which, int bits, int reps, double *avg_us, double *bytes) = 0;
    virtual int tune(int bits, int reps, int *chosen) = 0;
};

template <int KB>
struct BlockSolver : BlockBase {
    hipStream_t stream = nullptr;
    BlockKrylov<double, double, double, KB> b64;
    std::unique_ptr<BlockKrylov<float, float, float, KB>> b32;
    std::unique_ptr<BlockKrylov<float, float, __half, KB>> b16;
    double *cm = nullptr;      // column-major staging (n x KB)
    int *active = nullptr;     // device mask of refining columns
    SolveState *ost = nullptr, *host_ost = nullptr;
    int n = 0;
    bool graphs = true;
    ~BlockSolver()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        (void)hipFree(cm);
        (void)hipFree(active);
        (void)hipFree(ost);
        if (host_ost) (void)hipHostFree(host_ost);
    }

    int create(const K64 &k)
    {
        kb = KB;
        n = k.n;
        stream = k.stream;
        CM_TRY(b64.create(k, stream));
        CM_TRY(dev_alloc(&cm, (size_t)n * KB));
        CM_TRY(dev_alloc(&active, (size_t)KB));
        CM_HIP(hipMalloc((void **)&ost, KB * sizeof(SolveState)));
        CM_HIP(hipHostMalloc((void **)&host_ost, KB * sizeof(SolveState), hipHostMallocDefault));
        return CM_OK;
    }

    size_t bytes() const override
    {
        size_t b = b64.bytes() + (size_t)n * KB * sizeof(double);
        if (b32) b += b32->bytes();
        if (b16) b += b16->bytes();
        return b;
    }

    void set_stream(hipStream_t s) override
    {
        stream = s;
        b64.stream = s;
        if (b32) b32->stream = s;
        if (b16) b16->stream = s;
    }

    void set_graphs(bool on) override
    {
        graphs = on;
        b64.use_graph = on;
        if (b32) b32->use_graph = on;
        if (b16) b16->use_graph = on;
    }

    void drop_graphs() override
    {
        (void)hipStreamSynchronize(stream);
        b64.drop_graphs();
        if (b32) b32->drop_graphs();
        if (b16) b16->drop_graphs();
    }

    // long-row buffers sized for the matrices' current block product formats
    int ensure_workspace() override
    {
        CM_TRY(b64.ensure_racc());
        if (b32) CM_TRY(b32->ensure_racc());
        if (b16) CM_TRY(b16->ensure_racc());
        return CM_OK;
    }

    template <class KI, class BI>
    int ensure_inner(std::unique_ptr<BI> &bi, const KI *ki)
    {
        if (bi) return CM_OK;
        if (!ki) return fail(CM_ERR_ARG, "mixed precision not enabled (cm_system_enable_mixed)");
        auto q = std::make_unique<BI>();
        CM_TRY(q->create(*ki, stream));
        CM_TRY(q->ensure_racc());
        q->use_graph = graphs;
        bi = std::move(q);
        return CM_OK;
    }

    int set(int nc, const double *B, const double *X0) override
    {
        ncols = nc;
        b64.ncols = nc;
        CM_TRY(b64.ensure_racc());
        if (b32) CM_TRY(b32->ensure_racc());
        if (b16) CM_TRY(b16->ensure_racc());
        const int g = (int)std::min<size_t>(((size_t)n * KB + BS - 1) / BS, 8192);
        if (B) {  // B null: the resident right-hand sides are kept, only x is reset
            CM_HIP(hipMemcpyAsync(cm, B, (size_t)n * nc * sizeof(double), hipMemcpyHostToDevice, stream));
            k_interleave<double, KB><<<g, BS, 0, stream>>>(n, nc, cm, b64.b);
        }
        if (X0) {
            CM_HIP(hipMemcpyAsync(cm, X0, (size_t)n * nc * sizeof(double), hipMemcpyHostToDevice, stream));
            k_interleave<double, KB><<<g, BS, 0, stream>>>(n, nc, cm, b64.x);
        } else {
            CM_HIP(hipMemsetAsync(b64.x, 0, (size_t)n * KB * sizeof(double), stream));
        }
        CM_HIP(hipStreamSynchronize(stream));
        return CM_OK;
    }

    int get(double *X) override
    {
        const int g = (int)std::min<size_t>(((size_t)n * KB + BS - 1) / BS, 8192);
        k_deinterleave<KB><<<g, BS, 0, stream>>>(n, ncols, b64.x, cm);
        CM_HIP(hipMemcpyAsync(X, cm, (size_t)n * ncols * sizeof(double), hipMemcpyDeviceToHost, stream));
        CM_HIP(hipStreamSynchronize(stream));
        return CM_OK;
    }

    int solve64(double tol, int NMAX, double *resval, int *iter, int *info) override
    {
        CM_TRY(b64.begin(tol, NMAX, nullptr));
        b64.setup_from_x();
        SolveState fin[KB];
        CM_TRY(b64.run(NMAX, fin));
        for (int j = 0; j < ncols; j++) {
            resval[j] = fin[j].rc / fin[j].ro;
            iter[j] = fin[j].iter;
            if (NMAX >= 0) info[j] = fin[j].info;  // the reference leaves info untouched when NMAX < 0
        }
        return CM_OK;
    }

    // Per-column FP64 refinement (the single-system rules of solve_mixed,
    // applied to each column): a column leaves the refinement when it
    // converges, stalls, turns non-finite or reaches max_outer; the inner
    // block solve runs only the columns still refining.
    template <class BI>
    int mixed_with(BI &bi, double tol, double inner_tol, int NMAX, int max_outer, double *resval, int *outer,
                   int *inner_total, int *info)
    {
        bi.ncols = ncols;
        CM_TRY(b64.ensure_racc());  // the FP64 residual products below
        std::vector<double> res0(ncols, -1.0), prev(ncols, -1.0);
        std::vector<int> act(KB, 0), fin_info(ncols, -1), loose(ncols, 0);
        int steps = 0, inner = 0;
        const int g = b64.vgrid();
        for (int o = 0;; o++) {
            for (int j = 0; j < KB; j++) {
                std::memset(&host_ost[j], 0, sizeof(SolveState));
                host_ost[j].done = (j < ncols && fin_info[j] < 0) ? 0 : 1;
            }
            CM_HIP(hipMemcpyAsync(ost, host_ost, KB * sizeof(SolveState), hipMemcpyHostToDevice, stream));
            b64.hfl[0].all_done = 0;
            b64.hfl[0].ncols = ncols;
            CM_HIP(hipMemcpyAsync(b64.fl, b64.hfl, sizeof(BlockFlags), hipMemcpyHostToDevice, stream));
            b64.spmm(*b64.A, b64.bA, b64.x, BEpiResid{b64.b, bi.b}, ost, false);
            CM_TRY(b64.check_racc());
            CM_HIP(hipMemcpyAsync(host_ost, ost, KB * sizeof(SolveState), hipMemcpyDeviceToHost, stream));
            CM_HIP(hipStreamSynchronize(stream));
            bool any = false;
            for (int j = 0; j < ncols; j++) {
                act[j] = 0;
                if (fin_info[j] >= 0) continue;
                const double rn = host_ost[j].res_outer;
                if (res0[j] < 0.0) res0[j] = rn;
                resval[j] = res0[j] > 0.0 ? rn / res0[j] : 0.0;
                if (!std::isfinite(rn)) fin_info[j] = 1;
                else if (!(rn > tol * res0[j])) fin_info[j] = 0;
                else if (o >= max_outer) fin_info[j] = 3;
                else if (prev[j] >= 0.0 && !(rn < prev[j]) && !loose[j]) fin_info[j] = 2;
                else {
                    act[j] = (prev[j] >= 0.0 && !(rn < prev[j])) ? 2 : 1;  // 2: retry at inner_tol
                    prev[j] = rn;
                    any = true;
                }
            }
            steps = o;
            if (!any) break;
            std::vector<double> itol(ncols, inner_tol);
            for (int j = 0; j < ncols; j++) {
                if (act[j] == 1) itol[j] = adaptive_inner_tol(inner_tol, tol, res0[j], host_ost[j].res_outer);
                loose[j] = act[j] && itol[j] > inner_tol;
                if (act[j]) act[j] = 1;
            }
            CM_TRY(bi.begin(inner_tol, NMAX, act.data(), itol.data()));
            bi.setup_zero();
            SolveState fin[KB];
            CM_TRY(bi.run(NMAX, fin));
            int most = 0;
            for (int j = 0; j < ncols; j++)
                if (act[j]) most = std::max(most, fin[j].ntrace);
            inner += most;
            CM_HIP(hipMemcpyAsync(active, act.data(), KB * sizeof(int), hipMemcpyHostToDevice, stream));
            k_block_x_accum<KB><<<g, BS, 0, stream>>>((size_t)n * KB, b64.x, bi.x, active);
            CM_HIP(hipGetLastError());
        }
        for (int j = 0; j < ncols; j++) info[j] = fin_info[j];
        if (outer) *outer = steps;
        if (inner_total) *inner_total = inner;
        return CM_OK;
    }

    int tune(int bits, int reps, int *chosen) override
    {
        if (reps < 1) return fail(CM_ERR_ARG, "reps must be positive");
        if (bits == 64) return b64.tune(reps, chosen);
        if (bits == 32) {
            CM_TRY(ensure_inner(b32, k32));
            return b32->tune(reps, chosen);
        }
        if (bits == 16) {
            CM_TRY(ensure_inner(b16, k16));
            return b16->tune(reps, chosen);
        }
        return fail(CM_ERR_ARG, "bits must be 64, 32 or 16");
    }

    int time_spmm(char which, int bits, int reps, double *avg_us, double *bytes) override
    {
        if (reps < 1) return fail(CM_ERR_ARG, "reps must be positive");
        if (bits == 64) return b64.time_spmm(which, reps, avg_us, bytes);
        if (bits == 32) {
            CM_TRY(ensure_inner(b32, k32));
            return b32->time_spmm(which, reps, avg_us, bytes);
        }
        if (bits == 16) {
            CM_TRY(ensure_inner(b16, k16));
            return b16->time_spmm(which, reps, avg_us, bytes);
        }
        return fail(CM_ERR_ARG, "bits must be 64, 32 or 16");
    }

    int solve_mixed(int bits, double tol, double inner_tol, int NMAX, int max_outer, double *resval, int *outer,
                    int *inner_total, int *info) override
    {
        if (bits == 32) {
            CM_TRY(ensure_inner(b32, k32));
            return mixed_with(*b32, tol, inner_tol, NMAX, max_outer, resval, outer, inner_total, info);
        }
        if (bits == 16) {
            CM_TRY(ensure_inner(b16, k16));
            return mixed_with(*b16, tol, inner_tol, NMAX, max_outer, resval, outer, inner_total, info);
        }
        return fail(CM_ERR_ARG, "factor_bits must be 32 or 16");
    }
};

static int block_width(int k)
{
    int kb = 1;
    while (kb < k) kb *= 2;
    return kb;
}

static int block_index(int kb)
{
    int i = 0;
    while ((1 << i) < kb) i++;
    return i;
}

// Device bytes a block solver of width kb would take on system k64 (FP64
// part; the inner solver of the given factor bits adds its own).
static size_t block_bytes_estimate(const K64 &k64, int kb, int bits)
{
    const int n = k64.n;
    auto slots = [&](int w) {
        size_t s = 0;
        for (const auto *m : {&k64.A, k64.sym ? &k64.GT : &k64.H, &k64.G}) {
            if (!m->present()) continue;
            // all rows short: n / (BS / w) workgroups bound the bins from above
            s = std::max(s, (size_t)(m->n + BS / w - 1) / (BS / w) + (size_t)(m->n + BS / WAVE - 1) / (BS / WAVE));
        }
        return s;
    };
    const size_t sl = slots(kb);
    // long-row remainder sums of the matrices on the diagonal block products
    size_t racc = 0;
    for (const auto *m : {&k64.A, k64.sym ? &k64.GT : &k64.H, &k64.G})
        if (m->block_dia && m->dia && m->dia->nlong)
            racc = std::max(racc, (size_t)m->dia->nlong * kb * (m->dia->ntile > 1 ? 1 + (size_t)m->dia->ntile : 1));
    auto one = [&](size_t vsz) {
        return 11 * (size_t)n * kb * vsz + std::max(sl, (size_t)MAX_RED_BLOCKS) * 4 * kb * sizeof(double) +
               3 * (size_t)n * sizeof(int) + kb * sizeof(SolveState) + sizeof(BlockFlags) + racc * vsz;
    };
    size_t b = one(sizeof(double)) + (size_t)n * kb * sizeof(double);
    if (bits == 32 || bits == 16) b += one(sizeof(float));
    return b;
}

static int make_block(std::unique_ptr<BlockBase> (&cache)[5], BlockBase *&cur, int k, const K64 &k64)
{
    if (k < 1 || k > 16) return fail(CM_ERR_ARG, "block width must be 1..16");
    if (k64.variant == 2 && k64.mode != 'E') return fail(CM_ERR_UNSUPPORTED, "block solves use apply mode E");
    const int kb = block_width(k);
    std::unique_ptr<BlockBase> &out = cache[block_index(kb)];
    if (!out) {
        auto mk = [&](auto tag) -> int {
            constexpr int KBV = decltype(tag)::value;
            auto b = std::make_unique<BlockSolver<KBV>>();
            CM_TRY(b->create(k64));
            out = std::move(b);
            return CM_OK;
        };
        int rc;
        switch (kb) {
        case 1: rc = mk(std::integral_constant<int, 1>{}); break;
        case 2: rc = mk(std::integral_constant<int, 2>{}); break;
        case 4: rc = mk(std::integral_constant<int, 4>{}); break;
        case 8: rc = mk(std::integral_constant<int, 8>{}); break;
        default: rc = mk(std::integral_constant<int, 16>{}); break;
        }
        if (rc != CM_OK) return rc;
    }
    cur = out.get();
    return CM_OK;
}

// Block products in the diagonal format: a matrix qualifies when its
// diagonal index exists and leaves at most BLOCK_DIA_REM of the nonzeros to
// the CSR remainder outside the long rows (remainders over SHORT_MAX entries,
// which a wave-per-row pass sums first).
constexpr double BLOCK_DIA_REM = 0.3;

// Results of the last mixed-precision solve.
struct MixedStats {
    int outer = 0, inner_total = 0, info = 0;
    double resval = 0.0;
};

struct System {
    hipStream_t stream = nullptr;
    int n = 0, variant = 0;
    std::unique_ptr<K64> k64;
    std::unique_ptr<K32> k32;
    std::unique_ptr<K16> k16;
    SolveState *ost = nullptr;       // device state of the refinement residual
    SolveState *host_ost = nullptr;  // pinned
    MixedStats mixed;
    // multi-RHS solvers, one per block width 1, 2, 4, 8, 16 (created on
    // first use; they share the matrices above); blk is the current one
    std::unique_ptr<BlockBase> blks[5];
    BlockBase *blk = nullptr;
    bool block_graphs = true;
    bool block_dia = true;  // block products in the diagonal format where a matrix suits it
    bool graphs = true;  // single-RHS passes from captured graphs
    size_t bytes() const
    {
        size_t b = k64->bytes() + (k32 ? k32->bytes() : 0) + (k16 ? k16->bytes() : 0);
        for (const auto &q : blks)
            if (q) b += q->bytes();
        return b;
    }
    ~System()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        blk = nullptr;
        for (auto &q : blks) q.reset();
        k16.reset();
        k32.reset();
        k64.reset();
        (void)hipFree(ost);
        if (host_ost) (void)hipHostFree(host_ost);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// Decides, per SpMV matrix (A, G, H, G^T) and precision copy, whether block
// products use the diagonal format (S.block_dia and at most BLOCK_DIA_REM of
// the nonzeros in the remainder), building the diagonal index and values on
// first use.  Idempotent; ends with the system stream synchronized, so block
// solves on other streams see the built arrays.
static int prepare_block_dia(System &S)
{
    K64 &k = *S.k64;
    DevCsr<double> *ms[] = {&k.A, &k.G, &k.H, &k.GT};
    for (int q = 0; q < 4; q++) {
        DevCsr<double> &m = *ms[q];
        if (!m.present()) continue;
        bool ok = false;
        if (S.block_dia) {
            CM_TRY(build_dia(m, k.stream));
            // the long remainders go to the wave-per-row kernel: the limit is on the others
            ok = m.dia && (double)(m.dia->nrem - m.dia->nlrem) <= BLOCK_DIA_REM * (double)m.nnz;
            if (ok) CM_TRY(build_dia_vals(m, k.stream));
        }
        m.block_dia = ok;
        auto low = [&](auto &c) -> int {
            follow_blocks(c, m);
            if (ok) CM_TRY(build_dia_vals(c, k.stream));
            c.block_dia = ok;
            return CM_OK;
        };
        if (S.k32) {
            DevCsr<float> *c32[] = {&S.k32->A, &S.k32->G, &S.k32->H, &S.k32->GT};
            CM_TRY(low(*c32[q]));
        }
        if (S.k16) {
            if (q == 0) CM_TRY(low(S.k16->A));
            if (q == 1) CM_TRY(low(S.k16->G));
            if (q == 2) CM_TRY(low(S.k16->H));
            if (q == 3) CM_TRY(low(S.k16->GT));
        }
    }
    CM_HIP(hipStreamSynchronize(k.stream));
    return CM_OK;
}

static int validate(const HostSys &h)
{
    if (h.n <= 0 || !h.Ai || !h.Aj || !h.Av || !h.IDv) return fail(CM_ERR_ARG, "null or empty system");
    if (h.variant < 0 || h.variant > 2) return fail(CM_ERR_ARG, "bad variant");
    const char mode = h.variant == 2 ? h.mode : 'E';
    if (mode != 'E' && mode != 'C' && mode != 'I' && mode != 'J') return fail(CM_ERR_ARG, "unknown apply mode");
    const bool sym = h.variant == 1 && h.a_symmetric;
    const bool need_gh = mode == 'E' || mode == 'C';
    if (!h.Gi || !h.Gj || !h.Gv) {
        if (need_gh) return fail(CM_ERR_ARG, "G factor missing");
    }
    if (need_gh && !sym && (!h.Hi || !h.Hj || !h.Hv)) return fail(CM_ERR_ARG, "H factor missing");
    const bool need_lu = h.variant == 2 && ((mode == 'C' && h.nr > 1) || mode == 'I' || mode == 'J');
    if (need_lu && (!h.Li || !h.Lj || !h.Lv || !h.Ui || !h.Uj || !h.Uv))
        return fail(CM_ERR_ARG, "L/U factors missing for this apply mode");
    return CM_OK;
}

static int system_create(System &S, const HostSys &h)
{
    CM_TRY(validate(h));
    S.n = h.n;
    S.variant = h.variant;
    CM_HIP(hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking));
    S.k64 = std::make_unique<K64>();
    return create_k64(*S.k64, h, S.stream);
}

static void print_verbose(int variant, int iter, int info, double rel)
{
    if (info == 0) {
        std::cout << "Convergence in " << iter << " iterations with relative residual " << rel << "\n";
    } else if (info == 1) {
        std::cout << "BiCGSTAB diverged and iterations stopped with relative residual " << rel << " in iteration "
                  << iter << "\n";
        std::cout << "Try to increase the quality of preconditioning\n";
    } else if (info == 2) {
        if (variant == 1) {
            std::cout << "BiCGSTAB diverged and iterations stopped with relative residual " << rel
                      << " in iteration " << iter << "\n";
            std::cout << "Try to increase the quality of preconditioning\n";
        } else {
            std::cout << "BiCGSTAB stagnated with relative residual " << rel << " in iteration " << iter << "\n";
            std::cout << "Try increasing prescribed tolerance\n";
        }
    } else if (info == 3) {
        std::cout << "Convergence failed after " << iter << " iterations with relative residual " << rel << "\n";
    }
}

// FP64 solve on the resident b and x.
static int solve64(System &S, double tol, int NMAX, double *resval, int *iter, int *info, double *trace,
                   int trace_cap, int *ntrace)
{
    K64 &k = *S.k64;
    CM_TRY(k.begin(tol, NMAX, info ? *info : 0, trace ? trace_cap : 0));
    k.setup_from_x();
    SolveState fin;
    CM_TRY(k.run(NMAX, fin));
    *iter = fin.iter;
    if (info) *info = fin.info;
    *resval = fin.rc / fin.ro;
    if (ntrace) *ntrace = fin.ntrace;
    if (trace && trace_cap > 0) {
        const int cnt = std::min(fin.ntrace, trace_cap);
        if (cnt > 0) CM_HIP(hipMemcpy(trace, k.trace, (size_t)cnt * sizeof(double), hipMemcpyDeviceToHost));
    }
    return CM_OK;
}

// Mixed precision: FP64 refinement around the low-precision inner solve.
//   r = b - A x (FP64 A, FP64 vectors), rn = ||r||; stop when rn <= tol * rn_0;
//   solve A d = r with the inner BiCGSTAB to inner_tol (FP32 A, FP32/FP16
//   factors, FP32 vectors, from d = 0); x += d in FP64.
// info: 0 converged; 2 an outer step did not reduce ||r|| (refinement stalled);
// 1 a non-finite residual; 3 max_outer steps without convergence.
template <class KI>
static int solve_mixed(System &S, KI &ki, double tol, double inner_tol, int NMAX, int max_outer, MixedStats &out)
{
    K64 &k = *S.k64;
    const int n = S.n;
    const int g = vec_grid(n);
    out = MixedStats{};
    double res0 = -1.0, prev = -1.0;
    bool loose = false;  // the last step ran at a tolerance looser than inner_tol
    for (int outer = 0;; outer++) {
        std::memset(S.host_ost, 0, sizeof(SolveState));
        CM_HIP(hipMemcpyAsync(S.ost, S.host_ost, sizeof(SolveState), hipMemcpyHostToDevice, S.stream));
        k.spmv(k.A, k.x, EpiResid{k.b, ki.b}, true, KC_SETUP, S.ost);
        CM_HIP(hipMemcpyAsync(S.host_ost, S.ost, sizeof(SolveState), hipMemcpyDeviceToHost, S.stream));
        CM_HIP(hipStreamSynchronize(S.stream));
        const double rn = S.host_ost->res_outer;
        if (res0 < 0.0) res0 = rn;
        out.resval = res0 > 0.0 ? rn / res0 : 0.0;
        out.outer = outer;
        if (!std::isfinite(rn)) {
            out.info = 1;
            return CM_OK;
        }
        if (!(rn > tol * res0)) {
            out.info = 0;
            return CM_OK;
        }
        if (outer >= max_outer) {
            out.info = 3;
            return CM_OK;
        }
        const bool stalled = prev >= 0.0 && !(rn < prev);
        if (stalled && !loose) {
            out.info = 2;
            return CM_OK;
        }
        prev = rn;
        const double itol = stalled ? inner_tol : adaptive_inner_tol(inner_tol, tol, res0, rn);
        loose = itol > inner_tol;
        CM_TRY(ki.begin(itol, NMAX, 0, 0));
        ki.setup_zero();
        SolveState fin;
        CM_TRY(ki.run(NMAX, fin));
        out.inner_total += fin.ntrace;
        k_x_accum<<<g, BS, 0, S.stream>>>(n, k.x, ki.x);
        CM_HIP(hipGetLastError());
    }
}

static int dropin_solve(System &S, const double *b, double *x, double tol, int NMAX, double *resval, int *iter,
                        int verbose, int *info)
{
    const int n = S.n;
    K64 &k = *S.k64;
    CM_HIP(hipMemcpyAsync(k.b, b, (size_t)n * sizeof(double), hipMemcpyHostToDevice, S.stream));
    CM_HIP(hipMemcpyAsync(k.x, x, (size_t)n * sizeof(double), hipMemcpyHostToDevice, S.stream));
    CM_TRY(solve64(S, tol, NMAX, resval, iter, info, nullptr, 0, nullptr));
    CM_HIP(hipMemcpy(x, k.x, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    // the reference prints only when verbose (AFIIM/src/bicgstab.cpp:227-281)
    if (verbose) print_verbose(S.variant, *iter, *info, *resval);
    return CM_OK;
}

}  // namespace cm

using namespace cm;

// ------------------------------------------------- persistent batched solves

// Host side of k_batch_solve: device copies of the system descriptors and
// items, the per-workgroup arena and the items' b / x, grown on demand and
// kept between calls.
struct BatchRunner {
    BatchSys *dsys = nullptr;
    BatchItem *ditems = nullptr, *hitems = nullptr;  // hitems pinned
    size_t sys_cap = 0, item_cap = 0;
    double *arena = nullptr, *vec = nullptr;
    size_t arena_cap = 0, vec_cap = 0;  // doubles
    int *head = nullptr;
    ~BatchRunner()
    {
        (void)hipFree(dsys);
        (void)hipFree(ditems);
        if (hitems) (void)hipHostFree(hitems);
        (void)hipFree(arena);
        (void)hipFree(vec);
        (void)hipFree(head);
    }
    size_t bytes() const
    {
        return sys_cap * sizeof(BatchSys) + item_cap * sizeof(BatchItem) + (arena_cap + vec_cap) * sizeof(double);
    }

    template <class P>
    static int grow(P **d, size_t &cap, size_t need)
    {
        if (need <= cap) return CM_OK;
        (void)hipFree(*d);
        *d = nullptr;
        CM_HIP(hipMalloc((void **)d, need * sizeof(P)));
        cap = need;
        return CM_OK;
    }

    // Solves nitems right-hand sides: item i is column B[i] (n of system
    // sys_of[i]) from initial guess X[i], result into X[i].
    int run(const std::vector<const System *> &systems, int nitems, const int *sys_of, const double *const *B,
            double *const *X, const double *tol, const int *nmax, int restart_true, double *resval, int *iter,
            int *info, int *cycles, hipStream_t s)
    {
        if (nitems <= 0) return CM_OK;
        const int ns = (int)systems.size();
        std::vector<BatchSys> hs(ns);
        int maxn = 0;
        for (int q = 0; q < ns; q++) {
            const K64 &k = *systems[q]->k64;
            if (k.variant == 2 && k.mode != 'E') return fail(CM_ERR_UNSUPPORTED, "batched solves use apply mode E");
            const DevCsr<double> &H = k.sym ? k.GT : k.H;
            hs[q] = BatchSys{k.n, k.variant, max_div_steps(k.variant, k.n), 0, k.A.ptr, k.A.col, k.A.val,
                             k.G.ptr, k.G.col, k.G.val, H.ptr, H.col, H.val, k.idv};
            maxn = std::max(maxn, k.n);
        }
        // largest systems first (longest solves start early)
        std::vector<int> order(nitems);
        for (int i = 0; i < nitems; i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return systems[sys_of[a]]->n > systems[sys_of[b]]->n; });
        std::vector<size_t> off(nitems);
        size_t tot = 0;
        for (int i = 0; i < nitems; i++) {
            off[i] = tot;
            tot += 2 * (size_t)systems[sys_of[i]]->n;
        }
        int dev = 0, cus = 256;
        CM_HIP(hipGetDevice(&dev));
        CM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        const int grid = std::min(nitems, cus);  // one PBS-thread workgroup fits a CU
        const size_t stride = ((size_t)9 * maxn + 63) / 64 * 64;
        CM_TRY(grow(&dsys, sys_cap, (size_t)ns));
        CM_TRY(grow(&ditems, item_cap, (size_t)nitems));
        if (nitems > (int)hitems_cap) {
            if (hitems) (void)hipHostFree(hitems);
            hitems = nullptr;
            CM_HIP(hipHostMalloc((void **)&hitems, (size_t)nitems * sizeof(BatchItem), hipHostMallocDefault));
            hitems_cap = nitems;
        }
        CM_TRY(grow(&arena, arena_cap, stride * grid));
        CM_TRY(grow(&vec, vec_cap, tot));
        if (!head) CM_HIP(hipMalloc((void **)&head, sizeof(int)));
        for (int q = 0; q < nitems; q++) {
            const int i = order[q];
            const int n = systems[sys_of[i]]->n;
            double *db = vec + off[i], *dx = db + n;
            CM_HIP(hipMemcpyAsync(db, B[i], (size_t)n * sizeof(double), hipMemcpyHostToDevice, s));
            CM_HIP(hipMemcpyAsync(dx, X[i], (size_t)n * sizeof(double), hipMemcpyHostToDevice, s));
            hitems[q] = BatchItem{sys_of[i], nmax[i], tol[i], db, dx, 0.0, 0, 0, 0, 0};
        }
        CM_HIP(hipMemcpyAsync(dsys, hs.data(), ns * sizeof(BatchSys), hipMemcpyHostToDevice, s));
        CM_HIP(hipMemcpyAsync(ditems, hitems, (size_t)nitems * sizeof(BatchItem), hipMemcpyHostToDevice, s));
        k_zero_int<<<1, 1, 0, s>>>(head);
        k_batch_solve<<<grid, PBS, 0, s>>>(nitems, dsys, ditems, arena, stride, head, restart_true);
        CM_HIP(hipGetLastError());
        CM_HIP(hipMemcpyAsync(hitems, ditems, (size_t)nitems * sizeof(BatchItem), hipMemcpyDeviceToHost, s));
        for (int q = 0; q < nitems; q++) {
            const int i = order[q];
            const int n = systems[sys_of[i]]->n;
            CM_HIP(hipMemcpyAsync(X[i], vec + off[i] + n, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s));
        }
        CM_HIP(hipStreamSynchronize(s));
        for (int q = 0; q < nitems; q++) {
            const int i = order[q];
            if (resval) resval[i] = hitems[q].resval;
            if (iter) iter[i] = hitems[q].iter;
            if (info && hitems[q].info >= 0) info[i] = hitems[q].info;  // untouched when NMAX < 0
            if (cycles) cycles[i] = hitems[q].cycles;
        }
        return CM_OK;
    }
    size_t hitems_cap = 0;
};

// Helpers of the GPU PDFIIM construction (cm_pdfiim_device).
namespace {
template <class T>
T *pd_new_copy(const std::vector<T> &v)
{
    T *p = new T[std::max<size_t>(v.size(), 1)];
    std::copy(v.begin(), v.end(), p);
    return p;
}

// Device buffers of one construction, freed on every exit.
struct PdBuffers {
    std::vector<void *> held;
    hipStream_t stream = nullptr;
    template <class T>
    int alloc(T **d, size_t count)
    {
        CM_TRY(dev_alloc(d, count));
        held.push_back((void *)*d);
        return CM_OK;
    }
    template <class T>
    int upload(T **d, const T *h, size_t count)
    {
        CM_TRY(alloc(d, count));
        if (count) CM_HIP(hipMemcpyAsync(*d, h, count * sizeof(T), hipMemcpyHostToDevice, stream));
        return CM_OK;
    }
    ~PdBuffers()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void *q : held) (void)hipFree(q);
        if (stream) (void)hipStreamDestroy(stream);
    }
};
}  // namespace

struct cm_system_s {
    System sys;
};

// The Krylov of one precision (64, 32, 16) and its matrix `which` (0..5: A,
// G, H, GT, L, U), or null.
template <class F>
static int with_matrix(System &S, int bits, int which, F f)
{
    if (which < 0 || which > 5) return fail(CM_ERR_ARG, "matrix index must be 0..5 (A, G, H, GT, L, U)");
    auto pick = [&](auto &ki) -> int {  // the matrices of one Krylov differ in value type
        switch (which) {
        case 0: return f(ki, ki.A);
        case 1: return f(ki, ki.G);
        case 2: return f(ki, ki.H);
        case 3: return f(ki, ki.GT);
        case 4: return f(ki, ki.L);
        default: return f(ki, ki.U);
        }
    };
    if (bits == 64) return pick(*S.k64);
    if (bits == 32 && S.k32) return pick(*S.k32);
    if (bits == 16 && S.k16) return pick(*S.k16);
    return fail(CM_ERR_ARG, "precision not enabled (64, or 32 / 16 after enabling mixed precision)");
}

extern "C" {

const char *cm_last_error(void) { return g_last_error.c_str(); }

int cm_device_count(void)
{
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

int cm_system_create(cm_system *out, int variant, int n, int a_symmetric, const int *Ai, const int *Aj,
                     const double *Av, const int *Gi, const int *Gj, const double *Gv, const double *IDv,
                     const int *Hi, const int *Hj, const double *Hv, const int *Li, const int *Lj,
                     const double *Lv, const int *Ui, const int *Uj, const double *Uv, int nr, char mode)
{
    if (!out) return fail(CM_ERR_ARG, "null handle");
    *out = nullptr;
    if (cm_device_count() <= 0) return fail(CM_ERR_NODEV, "no GPU visible");
    auto h = std::make_unique<cm_system_s>();
    const HostSys hs{variant, n, a_symmetric, nr, mode, Ai, Aj, Av, Gi, Gj, Gv, IDv, Hi, Hj, Hv, Li, Lj, Lv, Ui, Uj, Uv};
    CM_TRY(system_create(h->sys, hs));
    *out = h.release();
    return CM_OK;
}

int cm_system_destroy(cm_system sys)
{
    delete sys;
    return CM_OK;
}

int cm_system_set_rhs(cm_system h, const double *b)
{
    if (!h || !b) return fail(CM_ERR_ARG, "null argument");
    System &S = h->sys;
    CM_HIP(hipMemcpyAsync(S.k64->b, b, (size_t)S.n * sizeof(double), hipMemcpyHostToDevice, S.stream));
    CM_HIP(hipStreamSynchronize(S.stream));
    return CM_OK;
}

int cm_system_set_x(cm_system h, const double *x)
{
    if (!h) return fail(CM_ERR_ARG, "null handle");
    System &S = h->sys;
    if (x) CM_HIP(hipMemcpyAsync(S.k64->x, x, (size_t)S.n * sizeof(double), hipMemcpyHostToDevice, S.stream));
    else CM_HIP(hipMemsetAsync(S.k64->x, 0, (size_t)S.n * sizeof(double), S.stream));
    CM_HIP(hipStreamSynchronize(S.stream));
    return CM_OK;
}

int cm_system_get_x(cm_system h, double *x)
{
    if (!h || !x) return fail(CM_ERR_ARG, "null argument");
    System &S = h->sys;
    CM_HIP(hipStreamSynchronize(S.stream));
    CM_HIP(hipMemcpy(x, S.k64->x, (size_t)S.n * sizeof(double), hipMemcpyDeviceToHost));
    return CM_OK;
}

int cm_system_solve(cm_system h, double tol, int NMAX, double *resval, int *iter, int *info, double *trace,
                    int trace_cap, int *ntrace)
{
    if (!h || !resval || !iter) return fail(CM_ERR_ARG, "null argument");
    return solve64(h->sys, tol, NMAX, resval, iter, info, trace, trace_cap, ntrace);
}

int cm_system_enable_mixed(cm_system h, int factor_bits)
{
    if (!h) return fail(CM_ERR_ARG, "null handle");
    System &S = h->sys;
    if (factor_bits != 32 && factor_bits != 16) return fail(CM_ERR_ARG, "factor_bits must be 32 or 16");
    if (!S.ost) {
        CM_HIP(hipMalloc((void **)&S.ost, sizeof(SolveState)));
        CM_HIP(hipHostMalloc((void **)&S.host_ost, sizeof(SolveState), hipHostMallocDefault));
    }
    if (factor_bits == 32 && !S.k32) {
        auto k = std::make_unique<K32>();
        CM_TRY(create_inner(*k, *S.k64));
        k->use_graph = S.graphs;
        S.k32 = std::move(k);
    }
    if (factor_bits == 16 && !S.k16) {
        auto k = std::make_unique<K16>();
        CM_TRY(create_inner(*k, *S.k64));
        k->use_graph = S.graphs;
        S.k16 = std::move(k);
    }
    return CM_OK;
}

int cm_system_solve_mixed(cm_system h, int factor_bits, double tol, double inner_tol, int NMAX, int max_outer,
                          double *resval, int *outer, int *inner_total, int *info)
{
    if (!h || !resval) return fail(CM_ERR_ARG, "null argument");
    System &S = h->sys;
    MixedStats ms;
    if (factor_bits == 32) {
        if (!S.k32) return fail(CM_ERR_ARG, "FP32 inner solver not enabled (cm_system_enable_mixed)");
        CM_TRY(solve_mixed(S, *S.k32, tol, inner_tol, NMAX, max_outer, ms));
    } else if (factor_bits == 16) {
        if (!S.k16) return fail(CM_ERR_ARG, "FP16-factor inner solver not enabled (cm_system_enable_mixed)");
        CM_TRY(solve_mixed(S, *S.k16, tol, inner_tol, NMAX, max_outer, ms));
    } else {
        return fail(CM_ERR_ARG, "factor_bits must be 32 or 16");
    }
    S.mixed = ms;
    *resval = ms.resval;
    if (outer) *outer = ms.outer;
    if (inner_total) *inner_total = ms.inner_total;
    if (info) *info = ms.info;
    return CM_OK;
}

int cm_system_block_set(cm_system h, int k, const double *B, const double *X0)
{
    if (!h) return fail(CM_ERR_ARG, "null argument");
    System &S = h->sys;
    if (!B) {  // keep the resident block of right-hand sides, reset x (to X0, or zero)
        if (!S.blk || S.blk->ncols != k)
            return fail(CM_ERR_ARG, "no resident block of this width: pass the right-hand sides");
        return S.blk->set(k, nullptr, X0);
    }
    CM_TRY(prepare_block_dia(S));
    CM_TRY(make_block(S.blks, S.blk, k, *S.k64));
    S.blk->k32 = S.k32.get();  // mixed precision may have been enabled since
    S.blk->k16 = S.k16.get();
    S.blk->set_graphs(S.block_graphs);
    return S.blk->set(k, B, X0);
}

int cm_system_block_dia(cm_system h, int on)
{
    if (!h) return fail(CM_ERR_ARG, "null handle");
    System &S = h->sys;
    S.block_dia = on != 0;
    for (auto &q : S.blks)  // captured block passes carry the old kernels
        if (q) q->drop_graphs();
    CM_TRY(prepare_block_dia(S));
    for (auto &q : S.blks)  // a matrix may now need its long-row buffer
        if (q) CM_TRY(q->ensure_workspace());
    return CM_OK;
}

int cm_system_graphs(cm_system h, int on)
{
    if (!h) return fail(CM_ERR_ARG, "null handle");
    System &S = h->sys;
    S.k64->use_graph = on != 0;
    if (S.k32) S.k32->use_graph = on != 0;
    if (S.k16) S.k16->use_graph = on != 0;
    S.graphs = on != 0;
    return CM_OK;
}

int cm_system_block_graphs(cm_system h, int on)
{
    if (!h) return fail(CM_ERR_ARG, "null handle");
    h->sys.block_graphs = on != 0;
    for (auto &q : h->sys.blks)
        if (q) q->set_graphs(on != 0);
    return CM_OK;
}

int cm_system_block_time_spmm(cm_system h, char which, int bits, int reps, double *avg_us, double *bytes)
{
    if (!h || !avg_us || !bytes || !h->sys.blk) return fail(CM_ERR_ARG, "no block set");
    h->sys.blk->k32 = h->sys.k32.get();
    h->sys.blk->k16 = h->sys.k16.get();
    return h->sys.blk->time_spmm(which, bits, reps, avg_us, bytes);
}

int cm_system_autotune(cm_system h, int reps, int *tiles)
{
    if (!h || reps < 1) return fail(CM_ERR_ARG, "bad argument");
    System &S = h->sys;
    K64 &k = *S.k64;
    DevCsr<double> *ms[] = {&k.A, &k.G, &k.H, &k.GT, &k.L, &k.U};
    hipEvent_t e0, e1;
    CM_HIP(hipEventCreate(&e0));
    CM_HIP(hipEventCreate(&e1));
    int maxblk = 0;
    // staged row blocks: tile per FP64 matrix
    for (int q = 0; q < 6; q++) {
        DevCsr<double> &m = *ms[q];
        if (tiles) tiles[q] = 0;
        if (!m.present()) continue;
        CM_TRY(set_format(m, FMT_STAGED, k.stream));
        double best = 1e300;
        int best_t = TILE;
        for (int tile : {1024, 2048, 4096}) {
            CM_TRY(retile(m, tile, k.stream));
            CM_TRY(k.ensure_partials(m.nblk));
            float ms_ = 0.f;
            CM_TRY(time_product(k, m, reps, e0, e1, &ms_));
            if (ms_ < best) {
                best = ms_;
                best_t = tile;
            }
        }
        const bool had16 = m.d16 != nullptr;
        CM_TRY(retile(m, best_t, k.stream));
        if (had16) CM_TRY(build_idx16(m, k.stream));  // relative to the new row blocks
        maxblk = std::max(maxblk, m.nblk);
        if (tiles) tiles[q] = best_t;
    }
    // slot-major and diagonal indexes of the SpMV matrices (A, G, H, GT, and
    // L, U where mode C's sweeps multiply by them), then the format per
    // matrix and precision
    for (int q = 0; q < (k.lu_spmv() ? 6 : 4); q++) {
        CM_TRY(build_slots(*ms[q], k.stream));
        if (ms[q]->slots) maxblk = std::max(maxblk, ms[q]->slots->nb);
        CM_TRY(build_dia(*ms[q], k.stream));
        if (ms[q]->dia) maxblk = std::max(maxblk, (ms[q]->n + BS - 1) / BS);
    }
    k.drop_graphs();  // their launches carry the old row blocks
    CM_TRY(k.ensure_partials(maxblk));
    CM_TRY(tune_formats(k, reps, e0, e1));
    if (S.k32) {
        K32 &c = *S.k32;
        follow_blocks(c.A, k.A), follow_blocks(c.G, k.G), follow_blocks(c.H, k.H);
        follow_blocks(c.GT, k.GT), follow_blocks(c.L, k.L), follow_blocks(c.U, k.U);
        c.drop_graphs();
        CM_TRY(c.ensure_partials(maxblk));
        CM_TRY(tune_formats(c, reps, e0, e1));
    }
    if (S.k16) {
        K16 &c = *S.k16;
        follow_blocks(c.A, k.A), follow_blocks(c.G, k.G), follow_blocks(c.H, k.H);
        follow_blocks(c.GT, k.GT), follow_blocks(c.L, k.L), follow_blocks(c.U, k.U);
        c.drop_graphs();
        CM_TRY(c.ensure_partials(maxblk));
        CM_TRY(tune_formats(c, reps, e0, e1));
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    CM_HIP(hipGetLastError());
    return CM_OK;
}

int cm_system_spmv_format(cm_system h, int bits, int which, int format)
{
    if (!h) return fail(CM_ERR_ARG, "null handle");
    System &S = h->sys;
    K64 &k = *S.k64;
    DevCsr<double> *ms[] = {&k.A, &k.G, &k.H, &k.GT, &k.L, &k.U};
    if (which < 0 || which > 5) return fail(CM_ERR_ARG, "matrix index must be 0..5 (A, G, H, GT, L, U)");
    const bool need_slots = (format == FMT_SLOT || format == FMT_SLOT16) && !ms[which]->slots;
    const bool need_dia = (format == FMT_DIA || format == FMT_DIA2) && !ms[which]->dia;
    if (which < 4 && ms[which]->present() && (need_slots || need_dia)) {
        if (need_slots) CM_TRY(build_slots(*ms[which], k.stream));
        if (need_dia) CM_TRY(build_dia(*ms[which], k.stream));
        for (int bits : {32, 16})  // the low-precision copies share the new index
            if ((bits == 32 && S.k32) || (bits == 16 && S.k16))
                CM_TRY(with_matrix(S, bits, which, [&](auto &, auto &m) -> int {
                    follow_blocks(m, *ms[which]);
                    return CM_OK;
                }));
    }
    return with_matrix(S, bits, which, [&](auto &ki, auto &m) -> int {
        CM_TRY(set_format(m, format, ki.stream));
        if (m.use_slot) CM_TRY(ki.ensure_partials(m.slots->nb));
        if (m.use_dia) CM_TRY(ki.ensure_partials((m.n + BS - 1) / BS));
        ki.drop_graphs();
        CM_HIP(hipStreamSynchronize(ki.stream));
        return CM_OK;
    });
}

int cm_system_dia_info(cm_system h, int which, int *nd, int *offsets, long long *nrem)
{
    if (!h || !nd || !offsets || !nrem) return fail(CM_ERR_ARG, "null argument");
    if (which < 0 || which > 5) return fail(CM_ERR_ARG, "matrix index must be 0..5 (A, G, H, GT, L, U)");
    K64 &k = *h->sys.k64;
    DevCsr<double> *ms[] = {&k.A, &k.G, &k.H, &k.GT, &k.L, &k.U};
    DevCsr<double> &m = *ms[which];
    *nd = 0;
    *nrem = 0;
    if (!m.present()) return fail(CM_ERR_ARG, "matrix not resident");
    CM_TRY(build_dia(m, k.stream));
    if (!m.dia) return CM_OK;
    *nd = m.dia->nd;
    *nrem = m.dia->nrem;
    CM_HIP(hipMemcpy(offsets, m.dia->off, (size_t)m.dia->nd * sizeof(int), hipMemcpyDeviceToHost));
    return CM_OK;
}

int cm_system_spmv_formats(cm_system h, int *formats)
{
    if (!h || !formats) return fail(CM_ERR_ARG, "bad argument");
    System &S = h->sys;
    for (int p = 0; p < 3; p++)
        for (int q = 0; q < 6; q++) {
            formats[6 * p + q] = -1;
            const int bits = p == 0 ? 64 : p == 1 ? 32 : 16;
            if (bits == 32 && !S.k32) continue;
            if (bits == 16 && !S.k16) continue;
            (void)with_matrix(S, bits, q, [&](auto &, auto &m) -> int {
                formats[6 * p + q] = get_format(m);
                return CM_OK;
            });
        }
    return CM_OK;
}

int cm_system_block_autotune(cm_system h, int bits, int reps, int *short_max)
{
    if (!h || !short_max || !h->sys.blk) return fail(CM_ERR_ARG, "no block set");
    h->sys.blk->k32 = h->sys.k32.get();
    h->sys.blk->k16 = h->sys.k16.get();
    return h->sys.blk->tune(bits, reps, short_max);
}

int cm_batch_solve(int nitems, const cm_system *systems, const double *const *B, double *const *X,
                   const double *tol, const int *NMAX, int restart_true, double *resval, int *iter, int *info)
{
    if (nitems < 0 || (nitems > 0 && (!systems || !B || !X || !tol || !NMAX))) return fail(CM_ERR_ARG, "null argument");
    if (nitems == 0) return CM_OK;
    static thread_local std::unique_ptr<BatchRunner> runner;
    static thread_local hipStream_t stream = nullptr;
    if (!stream) CM_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    if (!runner) runner = std::make_unique<BatchRunner>();
    // distinct systems, in first-appearance order
    std::vector<const System *> sys;
    std::vector<int> sys_of(nitems);
    for (int i = 0; i < nitems; i++) {
        if (!systems[i] || !B[i] || !X[i]) return fail(CM_ERR_ARG, "null system or vector");
        const System *p = &systems[i]->sys;
        int q = 0;
        while (q < (int)sys.size() && sys[q] != p) q++;
        if (q == (int)sys.size()) sys.push_back(p);
        sys_of[i] = q;
    }
    return runner->run(sys, nitems, sys_of.data(), B, X, tol, NMAX, restart_true, resval, iter, info, nullptr,
                       stream);
}

// ---------------------------------------- batched solves without allocation
//
// cm_batch_solve_noalloc runs k_batch_solve with every buffer beyond the
// caller's inputs and outputs in one file-scope device array: the queue
// head, the system and item tables and the per-workgroup work arena.  Host
// code never passes the array's name to a kernel or a copy (in host code it
// names a host-side stand-in): the device address of each GPU's copy comes
// from hipGetSymbolAddress, taken once per device with that device current.
// Host-side tables are file-scope arrays as well, so a call allocates
// nothing, on the device or the host.  Calls are serialized by a lock; the
// array keeps the last call's contents (the queue head is cleared by a
// kernel at the start of each call).
namespace {
constexpr int NA_MAX_ITEMS = 4096, NA_MAX_SYS = 256;
constexpr size_t NA_ALIGN = 256;
constexpr size_t na_round(size_t b) { return (b + NA_ALIGN - 1) / NA_ALIGN * NA_ALIGN; }
constexpr size_t NA_HEAD = 0;
constexpr size_t NA_SYS = NA_ALIGN;
constexpr size_t NA_ITEMS = NA_SYS + na_round(NA_MAX_SYS * sizeof(BatchSys));
constexpr size_t NA_ARENA = NA_ITEMS + na_round(NA_MAX_ITEMS * sizeof(BatchItem));
constexpr size_t NA_ARENA_BYTES = (size_t)64 << 20;  // 64 MiB of work vectors
constexpr size_t NA_SCRATCH_BYTES = NA_ARENA + NA_ARENA_BYTES;
constexpr int NA_MAX_DEVICES = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[NA_SCRATCH_BYTES];

static std::mutex g_na_mu;
static unsigned char *g_na_scratch[NA_MAX_DEVICES];  // device address of `scratch` per device
static BatchSys g_na_sys[NA_MAX_SYS];
static BatchItem g_na_items[NA_MAX_ITEMS];
static const System *g_na_sysptr[NA_MAX_SYS];
static int g_na_sys_of[NA_MAX_ITEMS], g_na_order[NA_MAX_ITEMS];

int cm_batch_solve_noalloc(int nitems, const cm_system *systems, const double *const *dB, double *const *dX,
                           const double *tol, const int *NMAX, int restart_true, double *resval, int *iter,
                           int *info, void *stream)
{
    if (nitems < 0 || (nitems > 0 && (!systems || !dB || !dX || !tol || !NMAX)))
        return fail(CM_ERR_ARG, "null argument");
    if (nitems == 0) return CM_OK;
    if (nitems > NA_MAX_ITEMS)
        return fail(CM_ERR_ARG, "cm_batch_solve_noalloc: more than " + std::to_string(NA_MAX_ITEMS) + " items");
    std::lock_guard<std::mutex> lock(g_na_mu);
    int dev = 0, cus = 256;
    CM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NA_MAX_DEVICES) return fail(CM_ERR_ARG, "device index out of range");
    if (!g_na_scratch[dev]) {
        void *q = nullptr;
        CM_HIP(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)));
        g_na_scratch[dev] = static_cast<unsigned char *>(q);
    }
    unsigned char *base = g_na_scratch[dev];
    CM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    // distinct systems in first-appearance order
    int ns = 0, maxn = 0;
    for (int i = 0; i < nitems; i++) {
        if (!systems[i] || !dB[i] || !dX[i]) return fail(CM_ERR_ARG, "null system or vector");
        const System *p = &systems[i]->sys;
        int q = 0;
        while (q < ns && g_na_sysptr[q] != p) q++;
        if (q == ns) {
            if (ns == NA_MAX_SYS)
                return fail(CM_ERR_ARG, "cm_batch_solve_noalloc: more than " + std::to_string(NA_MAX_SYS) +
                                            " distinct systems");
            const K64 &k = *p->k64;
            if (k.variant == 2 && k.mode != 'E') return fail(CM_ERR_UNSUPPORTED, "batched solves use apply mode E");
            const DevCsr<double> &H = k.sym ? k.GT : k.H;
            g_na_sys[ns] = BatchSys{k.n, k.variant, max_div_steps(k.variant, k.n), 0, k.A.ptr, k.A.col, k.A.val,
                                    k.G.ptr, k.G.col, k.G.val, H.ptr, H.col, H.val, k.idv};
            g_na_sysptr[ns++] = p;
            maxn = std::max(maxn, k.n);
        }
        g_na_sys_of[i] = q;
    }
    const size_t stride = ((size_t)9 * maxn + 63) / 64 * 64;  // doubles per workgroup
    const size_t slots = NA_ARENA_BYTES / (stride * sizeof(double));
    if (slots < 1)
        return fail(CM_ERR_ARG, "cm_batch_solve_noalloc: a system of " + std::to_string(maxn) +
                                    " rows needs more than the " + std::to_string(NA_ARENA_BYTES >> 20) +
                                    " MiB work arena");
    const int grid = (int)std::min<size_t>({(size_t)nitems, (size_t)cus, slots});
    // largest systems first; ties in item order
    for (int i = 0; i < nitems; i++) g_na_order[i] = i;
    std::sort(g_na_order, g_na_order + nitems, [](int a, int b) {
        const int na = g_na_sysptr[g_na_sys_of[a]]->n, nb = g_na_sysptr[g_na_sys_of[b]]->n;
        return na != nb ? na > nb : a < b;
    });
    for (int q = 0; q < nitems; q++) {
        const int i = g_na_order[q];
        g_na_items[q] = BatchItem{g_na_sys_of[i], NMAX[i], tol[i], dB[i], dX[i], 0.0, 0, 0, 0, 0};
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    int *head = reinterpret_cast<int *>(base + NA_HEAD);
    BatchSys *dsys = reinterpret_cast<BatchSys *>(base + NA_SYS);
    BatchItem *ditems = reinterpret_cast<BatchItem *>(base + NA_ITEMS);
    double *arena = reinterpret_cast<double *>(base + NA_ARENA);
    CM_HIP(hipMemcpyAsync(dsys, g_na_sys, ns * sizeof(BatchSys), hipMemcpyHostToDevice, s));
    CM_HIP(hipMemcpyAsync(ditems, g_na_items, (size_t)nitems * sizeof(BatchItem), hipMemcpyHostToDevice, s));
    k_zero_int<<<1, 1, 0, s>>>(head);
    k_batch_solve<<<grid, PBS, 0, s>>>(nitems, dsys, ditems, arena, stride, head, restart_true);
    CM_HIP(hipGetLastError());
    CM_HIP(hipMemcpyAsync(g_na_items, ditems, (size_t)nitems * sizeof(BatchItem), hipMemcpyDeviceToHost, s));
    CM_HIP(hipStreamSynchronize(s));
    for (int q = 0; q < nitems; q++) {
        const int i = g_na_order[q];
        if (resval) resval[i] = g_na_items[q].resval;
        if (iter) iter[i] = g_na_items[q].iter;
        if (info && g_na_items[q].info >= 0) info[i] = g_na_items[q].info;  // untouched when NMAX < 0
    }
    return CM_OK;
}

// ------------------------------------------- PDFIIM construction on the GPU

int cm_pdfiim_device(int n, const double *Av, const int *Aj, const int *Ai, double **Gval, int **Gindj,
                     int **Gindi, double **IDv, double **Hval, int **Hindj, int **Hindi, int prefill, int lfill,
                     int postfill, double ftol, int warpsize, char filttype, char patt, double *device_ms)
{
    if (!Gval || !Gindj || !Gindi || !IDv || !Hval || !Hindj || !Hindi) return fail(CM_ERR_ARG, "null output");
    *Gval = *Hval = *IDv = nullptr;
    *Gindj = *Gindi = *Hindj = *Hindi = nullptr;
    if (n <= 0 || !Av || !Aj || !Ai) return fail(CM_ERR_ARG, "null or empty matrix");
    if (patt != 'B' && patt != 'L' && patt != 'U') return fail(CM_ERR_ARG, "patt must be 'B', 'L' or 'U'");
    // the reference's 'L'/'U' level-set expansion walks a stale list cursor
    // past the first pass (PDFIIM/src/parPrec.cpp:393-456): undefined there
    if (patt != 'B' && lfill > 1) return fail(CM_ERR_UNSUPPORTED, "patt 'L'/'U' need lfill == 1");
    // delete_last below an empty set is undefined in the reference
    if (warpsize != -1 && warpsize < 1) return fail(CM_ERR_ARG, "warpsize must be -1 or positive");
    CM_TRY(check_csr(n, Ai, Aj));
    std::vector<int> tp, tc;
    std::vector<double> tv;
    host_transpose(n, Ai, Aj, Av, tp, tc, tv);
    PdBuffers M;
    CM_HIP(hipStreamCreateWithFlags(&M.stream, hipStreamNonBlocking));
    const hipStream_t s = M.stream;
    const long long nnz = Ai[n];
    PdCsr A{}, At{};
    {
        int *p, *c, *tpd, *tcd;
        double *v, *tvd;
        CM_TRY(M.upload(&p, Ai, (size_t)n + 1));
        CM_TRY(M.upload(&c, Aj, (size_t)nnz));
        CM_TRY(M.upload(&v, Av, (size_t)nnz));
        CM_TRY(M.upload(&tpd, tp.data(), tp.size()));
        CM_TRY(M.upload(&tcd, tc.data(), tc.size()));
        CM_TRY(M.upload(&tvd, tv.data(), tv.size()));
        A = PdCsr{p, c, v};
        At = PdCsr{tpd, tcd, tvd};
    }
    const PdParams P{n, prefill, lfill, postfill, warpsize, ftol, std::max(1e-4, ftol), filttype, patt};
    int *gcnt, *hcnt, *flag;
    double *dinv;
    CM_TRY(M.alloc(&gcnt, (size_t)n));
    CM_TRY(M.alloc(&hcnt, (size_t)n));
    CM_TRY(M.alloc(&flag, (size_t)n));
    CM_TRY(M.alloc(&dinv, (size_t)n));
    hipEvent_t e0, e1;
    CM_HIP(hipEventCreate(&e0));
    CM_HIP(hipEventCreate(&e1));
    float ms_total = 0.f;
    // tier 1: every row 1..n-1 with patterns of at most 32 indices
    const int nrows = n - 1;
    constexpr int OUT1 = 33, OUT2 = 65;
    int *gi1 = nullptr, *hi1 = nullptr;
    double *gv1 = nullptr, *hv1 = nullptr;
    std::vector<int> hflag(n, 0), rows2, rows3;
    if (nrows > 0) {
        CM_TRY(M.alloc(&gi1, (size_t)nrows * OUT1));
        CM_TRY(M.alloc(&hi1, (size_t)nrows * OUT1));
        CM_TRY(M.alloc(&gv1, (size_t)nrows * OUT1));
        CM_TRY(M.alloc(&hv1, (size_t)nrows * OUT1));
        const PdOut O{gcnt, hcnt, flag, gi1, hi1, gv1, hv1, dinv, OUT1};
        const int grid = std::min(nrows, 16384);
        CM_HIP(hipEventRecord(e0, s));
        k_pdfiim_rows<32><<<grid, PD_WAVE, 0, s>>>(nrows, nullptr, P, A, At, O);
        CM_HIP(hipGetLastError());
        CM_HIP(hipEventRecord(e1, s));
        CM_HIP(hipMemcpyAsync(hflag.data() + 1, flag + 1, (size_t)nrows * sizeof(int), hipMemcpyDeviceToHost, s));
        CM_HIP(hipStreamSynchronize(s));
        float ms = 0.f;
        CM_HIP(hipEventElapsedTime(&ms, e0, e1));
        ms_total += ms;
        for (int i = 1; i < n; i++)
            if (hflag[i]) rows2.push_back(i);
    }
    // tier 2: the rows whose pattern outgrew 32 indices, with room for 64
    int *gi2 = nullptr, *hi2 = nullptr, *rows2d = nullptr;
    double *gv2 = nullptr, *hv2 = nullptr;
    const int n2 = (int)rows2.size();
    if (n2) {
        CM_TRY(M.upload(&rows2d, rows2.data(), rows2.size()));
        CM_TRY(M.alloc(&gi2, (size_t)n2 * OUT2));
        CM_TRY(M.alloc(&hi2, (size_t)n2 * OUT2));
        CM_TRY(M.alloc(&gv2, (size_t)n2 * OUT2));
        CM_TRY(M.alloc(&hv2, (size_t)n2 * OUT2));
        const PdOut O{gcnt, hcnt, flag, gi2, hi2, gv2, hv2, dinv, OUT2};
        CM_HIP(hipEventRecord(e0, s));
        k_pdfiim_rows<64><<<std::min(n2, 16384), PD_WAVE, 0, s>>>(n2, rows2d, P, A, At, O);
        CM_HIP(hipGetLastError());
        CM_HIP(hipEventRecord(e1, s));
        std::vector<int> f2(n);
        CM_HIP(hipMemcpyAsync(f2.data(), flag, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, s));
        CM_HIP(hipStreamSynchronize(s));
        float ms = 0.f;
        CM_HIP(hipEventElapsedTime(&ms, e0, e1));
        ms_total += ms;
        for (int i : rows2)
            if (f2[i]) rows3.push_back(i);
    }
    // tier 3: patterns of up to PD_BIG indices, the dense systems in global memory
    int *gi3 = nullptr, *hi3 = nullptr, *rows3d = nullptr;
    double *gv3 = nullptr, *hv3 = nullptr;
    const int n3 = (int)rows3.size();
    constexpr int OUT3 = PD_BIG + 1;
    if (n3) {
        const int grid3 = std::min(n3, 256);
        PdBigSlab *slabs;
        CM_TRY(M.alloc(&slabs, (size_t)grid3));
        CM_TRY(M.upload(&rows3d, rows3.data(), rows3.size()));
        CM_TRY(M.alloc(&gi3, (size_t)n3 * OUT3));
        CM_TRY(M.alloc(&hi3, (size_t)n3 * OUT3));
        CM_TRY(M.alloc(&gv3, (size_t)n3 * OUT3));
        CM_TRY(M.alloc(&hv3, (size_t)n3 * OUT3));
        const PdOut O{gcnt, hcnt, flag, gi3, hi3, gv3, hv3, dinv, OUT3};
        CM_HIP(hipEventRecord(e0, s));
        k_pdfiim_rows_big<<<grid3, PD_WAVE, 0, s>>>(n3, rows3d, P, A, At, O, slabs);
        CM_HIP(hipGetLastError());
        CM_HIP(hipEventRecord(e1, s));
        std::vector<int> f3(n);
        CM_HIP(hipMemcpyAsync(f3.data(), flag, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, s));
        CM_HIP(hipStreamSynchronize(s));
        float ms = 0.f;
        CM_HIP(hipEventElapsedTime(&ms, e0, e1));
        ms_total += ms;
        for (int i : rows3)
            if (f3[i]) {
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                return fail(CM_ERR_UNSUPPORTED, "PDFIIM pattern of row " + std::to_string(i) + " exceeds " +
                                                    std::to_string(PD_BIG) +
                                                    " indices; build these factors with cm_pdfiim");
            }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // counts and D back; packed offsets (row 0: the unit entry alone)
    std::vector<int> gc(n, 1), hc(n, 1);
    std::vector<double> dv(n);
    if (nrows > 0) {
        CM_HIP(hipMemcpyAsync(gc.data() + 1, gcnt + 1, (size_t)nrows * sizeof(int), hipMemcpyDeviceToHost, s));
        CM_HIP(hipMemcpyAsync(hc.data() + 1, hcnt + 1, (size_t)nrows * sizeof(int), hipMemcpyDeviceToHost, s));
        CM_HIP(hipMemcpyAsync(dv.data() + 1, dinv + 1, (size_t)nrows * sizeof(double), hipMemcpyDeviceToHost, s));
        CM_HIP(hipStreamSynchronize(s));
    }
    dv[0] = 1. / Av[0];  // PDFIIM/src/parPrec.cpp:132
    std::vector<long long> go(n + 1, 0), ho(n + 1, 0);
    for (int i = 0; i < n; i++) {
        go[i + 1] = go[i] + gc[i];
        ho[i + 1] = ho[i] + hc[i];
    }
    if (go[n] > INT_MAX || ho[n] > INT_MAX) return fail(CM_ERR_UNSUPPORTED, "factor has more than 2^31 entries");
    std::vector<int> gpk(go[n]), hpk(ho[n]);
    std::vector<double> gvk(go[n]), hvk(ho[n]);
    gpk[0] = 0;
    gvk[0] = 1.0;
    hpk[0] = 0;
    hvk[0] = 1.0;
    if (nrows > 0) {
        long long *god, *hod;
        int *gpd, *hpd;
        double *gvd, *hvd;
        CM_TRY(M.upload(&god, go.data(), go.size()));
        CM_TRY(M.upload(&hod, ho.data(), ho.size()));
        CM_TRY(M.alloc(&gpd, (size_t)go[n]));
        CM_TRY(M.alloc(&hpd, (size_t)ho[n]));
        CM_TRY(M.alloc(&gvd, (size_t)go[n]));
        CM_TRY(M.alloc(&hvd, (size_t)ho[n]));
        const int g1 = (int)std::min<long long>((nrows + BS - 1) / BS, 8192);
        k_pd_pack<<<g1, BS, 0, s>>>(nrows, nullptr, OUT1, gcnt, god, gi1, gv1, gpd, gvd);
        k_pd_pack<<<g1, BS, 0, s>>>(nrows, nullptr, OUT1, hcnt, hod, hi1, hv1, hpd, hvd);
        if (n2) {  // the tier-2 rows overwrite their (stale) tier-1 slots' packed places
            const int g2 = (int)std::min<long long>((n2 + BS - 1) / BS, 8192);
            k_pd_pack<<<g2, BS, 0, s>>>(n2, rows2d, OUT2, gcnt, god, gi2, gv2, gpd, gvd);
            k_pd_pack<<<g2, BS, 0, s>>>(n2, rows2d, OUT2, hcnt, hod, hi2, hv2, hpd, hvd);
        }
        if (n3) {  // and the tier-3 rows theirs
            const int g3 = (int)std::min<long long>((n3 + BS - 1) / BS, 8192);
            k_pd_pack<<<g3, BS, 0, s>>>(n3, rows3d, OUT3, gcnt, god, gi3, gv3, gpd, gvd);
            k_pd_pack<<<g3, BS, 0, s>>>(n3, rows3d, OUT3, hcnt, hod, hi3, hv3, hpd, hvd);
        }
        CM_HIP(hipGetLastError());
        const size_t g0 = (size_t)go[1], h0 = (size_t)ho[1];
        CM_HIP(hipMemcpyAsync(gpk.data() + g0, gpd + g0, (go[n] - g0) * sizeof(int), hipMemcpyDeviceToHost, s));
        CM_HIP(hipMemcpyAsync(gvk.data() + g0, gvd + g0, (go[n] - g0) * sizeof(double), hipMemcpyDeviceToHost, s));
        CM_HIP(hipMemcpyAsync(hpk.data() + h0, hpd + h0, (ho[n] - h0) * sizeof(int), hipMemcpyDeviceToHost, s));
        CM_HIP(hipMemcpyAsync(hvk.data() + h0, hvd + h0, (ho[n] - h0) * sizeof(double), hipMemcpyDeviceToHost, s));
        CM_HIP(hipStreamSynchronize(s));
    }
    // H rows are the packed row lists; G by rows from its column lists
    // (columns ascending in every row, the unit diagonal first)
    std::vector<int> hptr(n + 1), gptr(n + 1, 0);
    for (int i = 0; i <= n; i++) hptr[i] = (int)ho[i];
    for (long long k = 0; k < go[n]; k++) gptr[gpk[k] + 1]++;
    for (int i = 0; i < n; i++) gptr[i + 1] += gptr[i];
    std::vector<int> gj(go[n]);
    std::vector<double> gvv(go[n]);
    {
        std::vector<int> fill(gptr.begin(), gptr.end() - 1);
        for (int i = 0; i < n; i++)
            for (long long k = go[i]; k < go[i + 1]; k++) {
                const int q = fill[gpk[k]]++;
                gj[q] = i;
                gvv[q] = gvk[k];
            }
    }
    *Gindi = pd_new_copy(gptr);
    *Gindj = pd_new_copy(gj);
    *Gval = pd_new_copy(gvv);
    *Hindi = pd_new_copy(hptr);
    *Hindj = pd_new_copy(hpk);
    *Hval = pd_new_copy(hvk);
    *IDv = pd_new_copy(dv);
    if (device_ms) *device_ms = ms_total;
    return CM_OK;
}

int cm_system_block_release(cm_system h)
{
    if (!h) return fail(CM_ERR_ARG, "null handle");
    h->sys.blk = nullptr;
    for (auto &q : h->sys.blks) q.reset();
    return CM_OK;
}

int cm_system_block_get(cm_system h, double *X)
{
    if (!h || !X || !h->sys.blk) return fail(CM_ERR_ARG, "no block set");
    return h->sys.blk->get(X);
}

int cm_system_block_solve(cm_system h, double tol, int NMAX, double *resval, int *iter, int *info)
{
    if (!h || !resval || !iter || !info || !h->sys.blk) return fail(CM_ERR_ARG, "no block set");
    return h->sys.blk->solve64(tol, NMAX, resval, iter, info);
}

int cm_system_block_solve_mixed(cm_system h, int factor_bits, double tol, double inner_tol, int NMAX, int max_outer,
                                double *resval, int *outer, int *inner_total, int *info)
{
    if (!h || !resval || !info || !h->sys.blk) return fail(CM_ERR_ARG, "no block set");
    h->sys.blk->k32 = h->sys.k32.get();
    h->sys.blk->k16 = h->sys.k16.get();
    return h->sys.blk->solve_mixed(factor_bits, tol, inner_tol, NMAX, max_outer, resval, outer, inner_total, info);
}

int cm_system_spmv(cm_system h, char which, const double *x, double *y)
{
    if (!h || !x || !y) return fail(CM_ERR_ARG, "null argument");
    System &S = h->sys;
    K64 &k = *S.k64;
    const DevCsr<double> *m = nullptr;
    switch (which) {
    case 'A': m = &k.A; break;
    case 'G': m = &k.G; break;
    case 'H': m = k.sym ? &k.GT : &k.H; break;
    case 'L': m = &k.L; break;
    case 'U': m = &k.U; break;
    case 'T': m = &k.GT; break;
    default: break;
    }
    if (!m || !m->present()) return fail(CM_ERR_ARG, "matrix not present");
    CM_HIP(hipMemcpyAsync(k.t1, x, (size_t)S.n * sizeof(double), hipMemcpyHostToDevice, S.stream));
    const bool prof_on = k.prof.on;  // kernel-level calls are not part of a solve's timing
    k.prof.on = false;
    k.spmv(*m, k.t1, EpiStore<double>{k.t2}, false, KC_A);
    k.prof.on = prof_on;
    CM_HIP(hipMemcpyAsync(y, k.t2, (size_t)S.n * sizeof(double), hipMemcpyDeviceToHost, S.stream));
    CM_HIP(hipStreamSynchronize(S.stream));
    CM_HIP(hipGetLastError());
    return CM_OK;
}

int cm_system_apply(cm_system h, const double *v, double *out)
{
    if (!h || !v || !out) return fail(CM_ERR_ARG, "null argument");
    System &S = h->sys;
    K64 &k = *S.k64;
    CM_HIP(hipMemcpyAsync(k.s, v, (size_t)S.n * sizeof(double), hipMemcpyHostToDevice, S.stream));
    const bool prof_on = k.prof.on;
    k.prof.on = false;
    k.enqueue_apply(k.s, k.sc, false);
    k.prof.on = prof_on;
    CM_HIP(hipMemcpyAsync(out, k.sc, (size_t)S.n * sizeof(double), hipMemcpyDeviceToHost, S.stream));
    CM_HIP(hipStreamSynchronize(S.stream));
    CM_HIP(hipGetLastError());
    return CM_OK;
}

long long cm_system_nnz(cm_system h, char which)
{
    if (!h) return -1;
    K64 &k = *h->sys.k64;
    const DevCsr<double> *m = nullptr;
    switch (which) {
    case 'A': m = &k.A; break;
    case 'G': m = &k.G; break;
    case 'H': m = k.sym ? &k.GT : &k.H; break;
    case 'L': m = &k.L; break;
    case 'U': m = &k.U; break;
    case 'T': m = &k.GT; break;
    default: break;
    }
    return (m && m->present()) ? m->nnz : -1;
}

int cm_system_profile(cm_system h, int mask)
{
    return cm_system_profile_sampled(h, mask, 0);
}

int cm_system_profile_sampled(cm_system h, int mask, int passes)
{
    if (!h || passes < 0) return fail(CM_ERR_ARG, "bad argument");
    System &S = h->sys;
    Prof *ps[] = {&S.k64->prof, S.k32 ? &S.k32->prof : nullptr, S.k16 ? &S.k16->prof : nullptr};
    for (Prof *p : ps)
        if (p) {
            p->on = (mask != 0);
            p->mask = (unsigned)mask;
            p->sample = passes;
            p->sampled = 0;
            p->reset_stats();
        }
    return CM_OK;
}

int cm_system_kernel_stats(cm_system h, int category, long long *launches, double *total_ms, double *total_bytes)
{
    if (!h || category < 0 || category >= KC_N) return fail(CM_ERR_ARG, "bad category");
    System &S = h->sys;
    long long nl = 0;
    double ms = 0.0, by = 0.0;
    const Prof *ps[] = {&S.k64->prof, S.k32 ? &S.k32->prof : nullptr, S.k16 ? &S.k16->prof : nullptr};
    for (const Prof *p : ps)
        if (p) {
            nl += p->launches[category];
            ms += p->ms[category];
            by += p->bytes[category];
        }
    if (launches) *launches = nl;
    if (total_ms) *total_ms = ms;
    if (total_bytes) *total_bytes = by;
    return CM_OK;
}

int cm_system_kernel_format_bytes(cm_system h, int category, double *total_bytes)
{
    if (!h || category < 0 || category >= KC_N || !total_bytes) return fail(CM_ERR_ARG, "bad category or null output");
    System &S = h->sys;
    double by = 0.0;
    const Prof *ps[] = {&S.k64->prof, S.k32 ? &S.k32->prof : nullptr, S.k16 ? &S.k16->prof : nullptr};
    for (const Prof *p : ps)
        if (p) by += p->fbytes[category];
    *total_bytes = by;
    return CM_OK;
}

// Formats of a one-shot system without timing anything: the diagonal format
// (built on the device) for each SpMV matrix whose diagonals held by 60% of
// the rows leave at most ONESHOT_DIA_REM of its nonzeros to the remainder --
// the second variant from DIA2_MIN_ND diagonals up, where it times
// faster -- the staged row blocks otherwise, also for a matrix the device
// build leaves to the host (whose index would take longer than the solve).
constexpr double ONESHOT_DIA_REM = 0.3;
static int oneshot_formats(System &S)
{
    K64 &k = *S.k64;
    for (DevCsr<double> *m : {&k.A, &k.G, &k.H, &k.GT, &k.L, &k.U}) {
        if (!m->present()) continue;
        if ((m == &k.L || m == &k.U) && !k.lu_spmv()) continue;  // only triangular solves read them
        bool done = false;
        if (!m->dia) CM_TRY(build_dia_device(*m, k.stream, &done));
        if (!m->dia || (double)m->dia->nrem > ONESHOT_DIA_REM * (double)m->nnz) continue;
        CM_TRY(set_format(*m, m->dia->nd >= DIA2_MIN_ND ? FMT_DIA2 : FMT_DIA, k.stream));
        CM_TRY(k.ensure_partials((m->n + BS - 1) / BS));
    }
    return CM_OK;
}

// One-shot systems are freed off the caller's thread: the hipFree calls for
// a system's few dozen buffers take ~10 ms that the call need not wait for.
// One at a time: the next one-shot call (and library unload) first waits
// for the previous release, so at most one released system is held.
class Releaser {
  public:
    void hand(System *s)
    {
        std::lock_guard<std::mutex> lk(mu);
        wait_locked();
        int dev = 0;
        (void)hipGetDevice(&dev);
        th = std::thread([s, dev] {
            (void)hipSetDevice(dev);
            delete s;
        });
    }
    void wait()
    {
        std::lock_guard<std::mutex> lk(mu);
        wait_locked();
    }
    ~Releaser() { wait(); }

  private:
    std::mutex mu;
    std::thread th;
    void wait_locked()
    {
        if (th.joinable()) th.join();
    }
};
static Releaser g_oneshot_release;

static int dropin(int variant, int n, int a_symmetric, const int *Ai, const int *Aj, const double *Av,
                  const double *b, double tol, int NMAX, const int *Gi, const int *Gj, const double *Gv,
                  const double *IDv, const int *Hi, const int *Hj, const double *Hv, const int *Li, const int *Lj,
                  const double *Lv, const int *Ui, const int *Uj, const double *Uv, int nr, char mode, double *x,
                  double *resval, int *iter, int verbose, int *info)
{
    if (!b || !x || !resval || !iter || !info) return fail(CM_ERR_ARG, "null argument");
    if (cm_device_count() <= 0) return fail(CM_ERR_NODEV, "no GPU visible");
    g_oneshot_release.wait();
    std::unique_ptr<System> S(new System());
    const HostSys hs{variant, n, a_symmetric, nr, mode, Ai, Aj, Av, Gi, Gj, Gv, IDv, Hi, Hj, Hv, Li, Lj, Lv, Ui, Uj, Uv};
    int rc = system_create(*S, hs);
    if (rc == CM_OK) rc = oneshot_formats(*S);
    if (rc == CM_OK) rc = dropin_solve(*S, b, x, tol, NMAX, resval, iter, verbose, info);
    if (rc == CM_OK) {
        g_oneshot_release.hand(S.release());
    }
    return rc;  // on an error the system is freed here, with the error's message kept
}

int cm_afiim_epbicgstab(int n, double *Av, int *Aj, int *Ai, double *b, double tol, int NMAX, double *Gv, int *Gj,
                        int *Gi, double *IDv, double *Hv, int *Hj, int *Hi, double *x, double *resval, int *iter,
                        int verbose, int *info)
{
    return dropin(0, n, 0, Ai, Aj, Av, b, tol, NMAX, Gi, Gj, Gv, IDv, Hi, Hj, Hv, nullptr, nullptr, nullptr,
                  nullptr, nullptr, nullptr, 1, 'E', x, resval, iter, verbose, info);
}

int cm_pdfiim_epbicgstab(int n, int a_symmetric, const int *Ai, const int *Aj, const double *Av, const double *b,
                         double tol, int NMAX, const int *Gi, const int *Gj, const double *Gv, const double *IDv,
                         const int *Hi, const int *Hj, const double *Hv, double *x, double *resval, int *iter,
                         int verbose, int *info)
{
    return dropin(1, n, a_symmetric, Ai, Aj, Av, b, tol, NMAX, Gi, Gj, Gv, IDv, Hi, Hj, Hv, nullptr, nullptr,
                  nullptr, nullptr, nullptr, nullptr, 1, 'E', x, resval, iter, verbose, info);
}

int cm_acifim_pbicgstab(int n, double *Av, int *Aj, int *Ai, double *b, double tol, int NMAX, double *Gv, int *Gj,
                        int *Gi, double *IDv, double *Hv, int *Hj, int *Hi, double *Lv, int *Lj, int *Li,
                        double *Uv, int *Uj, int *Ui, int nr, char mode, double *x, double *resval, int *iter,
                        int verbose, int *info)
{
    return dropin(2, n, 0, Ai, Aj, Av, b, tol, NMAX, Gi, Gj, Gv, IDv, Hi, Hj, Hv, Li, Lj, Lv, Ui, Uj, Uv, nr, mode,
                  x, resval, iter, verbose, info);
}

}  // extern "C"

#include "cm_serve.h"
#include "cm_serve.hpp"
