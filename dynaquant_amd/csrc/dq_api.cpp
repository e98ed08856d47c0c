// dq_api.cpp -- host side of the C ABI (include/dynaquant.h).
//
// Owns all device memory of a system, drives the per-step kernel sequence
// on one HIP stream (optionally replayed from a captured hipGraph) and
// converts between the caller's [b][i] interleaved-complex buffers and the
// engine's [i][b] device layout.  The reference keeps this orchestration in
// propagate.cu:167-298 and systemClass.cu, with a host sync after every
// SpMV; here the host only enqueues.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <utility>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dynaquant.h"
#include "dq_internal.h"
#include "dq_adjoint.h"

using dq::Dev;

namespace {

thread_local std::string g_err;

dq_status fail(dq_status st, const std::string& msg)
{
    g_err = msg;
    return st;
}

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(e_ == hipErrorOutOfMemory ? DQ_ERR_OOM : DQ_ERR_HIP,            \
                        std::string(#expr) + ": " + hipGetErrorString(e_));             \
    } while (0)

#define DQ_TRY(expr)                       \
    do {                                   \
        dq_status s_ = (expr);             \
        if (s_ != DQ_OK) return s_;        \
    } while (0)

double reference_hbar()
{
    // cppUtility.cpp:19-22, same evaluation order
    double pi = std::atan(1) * 4;
    double lightspeed = 299792548.;
    return (1 / 100.) * (1e15) * 1. / (lightspeed * 2 * pi);
}

}  // namespace

namespace dq {
thread_local KernelEvents g_kev;
dq_status set_error(dq_status st, const std::string& msg) { return fail(st, msg); }
}

// Row-blocked copy of a coupling matrix for the batched SpMM: per block of R
// rows the sorted union of their columns and R values per union column (0
// where a row lacks that column).
struct RowBlocks {
    int32_t* ptr = nullptr;
    int32_t* col = nullptr;
    float* val32 = nullptr;
    double* val64 = nullptr;
    uint16_t* lo16 = nullptr;   // R = 16: bfloat16 low parts val - (double)val32, MFMA operand order
};

// LDS one workgroup of the ring SpMM may use: the device's opt-in limit
// (160 KiB on gfx950), queried once; 0 when unknown
static size_t ring_lds_limit()
{
    static size_t lim = [] {
        int dev = 0, a = 0, b = 0;
        if (hipGetDevice(&dev) != hipSuccess) return (size_t)0;
        if (hipDeviceGetAttribute(&a, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) a = 0;
        if (hipDeviceGetAttribute(&b, hipDeviceAttributeSharedMemPerBlockOptin, dev) != hipSuccess) b = 0;
        return (size_t)std::max(a, b);
    }();
    return lim;
}

// Lists of the ring SpMM (dq_ring.hip): per 4-row block its column union in
// even-aligned pairs, in items of up to 16 pairs: the couplings and their
// bfloat16 low parts in MFMA operand order and the LDS offsets of the pairs;
// per 128-row chunk the window of x rows it references.
struct RingLists {
    int32_t* item = nullptr;         // [ceil(n/4) + 1] first item of each block
    float* a = nullptr;              // [item][64][2]
    uint2* alo = nullptr;            // [item][32]
    uint4* off = nullptr;            // [item][4]
    int2* win = nullptr;             // [ceil(n/128)] [lo, hi)
};

struct dq_system {
    dq_config cfg{};
    Dev d{};
    hipStream_t stream = nullptr;
    static constexpr int kMaxSlices = 4;
    hipStream_t side[kMaxSlices - 1] = {};   // streams of slices 1.. in captured steps
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join[kMaxSlices - 1] = {};
    int slices = 1;                   // captured steps run this many trajectory slices concurrently
    RingLists ring;                   // ring SpMM lists (built while ring_rows > 0 and H is set)
    int ring_rows = 0;                // ring SpMM: rows per workgroup (0: the row-blocked SpMMs)
    // substep control (dq_system_set_krylov_control): chunks of kchunk steps,
    // K doubled up to kmax while K x the largest cycle estimate exceeds ktol
    double hbar = 0.0;
    double ktol = 0.0;
    int kmax = 0, kchunk = 10, K0 = 1;
    int64_t k_rejects = 0;
    double* ck_err_max = nullptr;     // [Bp] device maxima, kept while control is off
    struct {
        double2* psi = nullptr;
        double *de = nullptr, *err = nullptr, *nrm2 = nullptr;
        void* diag = nullptr;
        int* status = nullptr;
        int64_t* step = nullptr;
    } ck;
    int ring_waves = 16;              // waves per workgroup of the ring SpMM
    bool ring_lds_set = false;        // dynamic LDS limit raised for the ring SpMM
    hipGraphExec_t graph = nullptr;
    double last_ms = 0.0;
    int n = 0, B = 0, Bp = 0, m = 0, K = 1;
    bool have_h = false, have_ou = false, have_state = false, have_disorder = false;
    bool have_hsr = false, have_zz = false;
    bool noise_ready() const
    {
        switch (d.noise_model) {
        case DQ_NOISE_OU: return have_ou;
        case DQ_NOISE_HSR: return have_hsr;
        case DQ_NOISE_ZZ: return have_zz;
        default: return true;
        }
    }
    int64_t step = 0;                 // host mirror of the device counter
    int64_t obs_cap_rec = 0;          // records allocated
    std::vector<void*> allocs;        // everything owned
    // host copies
    std::vector<int32_t> rowptr, col;
    std::vector<double> val;
    // device column (slot) of caller trajectory b; slots are ordered by
    // n_active, largest first, so the batched kernels can skip rows that are
    // inactive for a whole slice of slots
    std::vector<int> slot;
    std::vector<int32_t> nact_slot;   // n_active per slot (0 for padding)
    // profiling: (class, start, stop) per bracketed launch of the current run
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<int> ev_class;
    size_t ev_used = 0;
    double kms[DQ_K_CLASSES] = {0};
    int64_t kcount[DQ_K_CLASSES] = {0};
    // observable config
    int ngroups = 0, every = 0, site_mode = 0;
    int64_t max_steps = 0;
    bool have_bra = false;
    // Taylor-Krylov propagation: order p, auxiliary Krylov m, the auxiliary
    // matrix (host CSR for the row blocks, device CSR, row blocks, diagonal
    // broadcast to [n][Bp] in the basis precision)
    struct {
        int p = 0, m = 0;
        std::vector<int32_t> rowptr_h, col_h;
        std::vector<double> val_h;
        int32_t *rowptr = nullptr, *col = nullptr;
        double* val64 = nullptr;
        float* val32 = nullptr;
        RowBlocks rb;
        void* diag = nullptr;
    } tk;

    ~dq_system()
    {
        if (graph) hipGraphExecDestroy(graph);
        for (hipEvent_t e : ev_pool) hipEventDestroy(e);
        for (void* p : allocs) hipFree(p);
        if (ev0) hipEventDestroy(ev0);
        if (ev1) hipEventDestroy(ev1);
        if (ev_fork) hipEventDestroy(ev_fork);
        for (int k = 0; k < kMaxSlices - 1; ++k) {
            if (ev_join[k]) hipEventDestroy(ev_join[k]);
            if (side[k]) hipStreamDestroy(side[k]);
        }
        if (stream) hipStreamDestroy(stream);
    }

    template <typename T>
    dq_status alloc(T** p, size_t count)
    {
        void* q = nullptr;
        const size_t bytes = count * sizeof(T) > 0 ? count * sizeof(T) : 16;
        if (getenv("DQ_TEST_FAIL_ALLOC")) return fail(DQ_ERR_OOM, "allocation refused (DQ_TEST_FAIL_ALLOC)");
        hipError_t e = hipMalloc(&q, bytes);
        if (e != hipSuccess) return fail(DQ_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
        allocs.push_back(q);
        hipMemsetAsync(q, 0, bytes, stream);
        *p = static_cast<T*>(q);
        return DQ_OK;
    }
    template <typename T>
    void release(T*& p)
    {
        if (!p) return;
        for (auto& q : allocs)
            if (q == (void*)p) { hipFree(q); q = nullptr; }
        p = nullptr;
    }
    void drop_graph()
    {
        if (graph) { hipGraphExecDestroy(graph); graph = nullptr; }
    }
};

extern "C" const char* dq_last_error(void) { return g_err.c_str(); }

// Device buffers a setter allocates before it replaces the old ones: released
// again unless keep() is called, so a setter that fails half-way leaves the
// system as it was.
struct Pending {
    dq_system* sys;
    std::vector<void*> ptrs;
    bool kept = false;
    explicit Pending(dq_system* s) : sys(s) {}
    template <typename T>
    dq_status alloc(T** p, size_t count)
    {
        dq_status st = sys->alloc(p, count);
        if (st == DQ_OK) ptrs.push_back((void*)*p);
        return st;
    }
    void keep() { kept = true; }
    ~Pending()
    {
        if (kept) return;
        for (void* p : ptrs) sys->release(p);
    }
};

// Slots of a sorted batch regrouped so that concurrent trajectory slices
// (contiguous runs of whole 64-slot groups) carry similar site counts: the
// 64-slot groups g0 >= g1 >= ... are dealt over four bins in snake order
// (g0 g1 g2 g3 g3' g2' ...: bin 0 takes g0, g7, g8, g15, ...), and the bins are
// laid out as 0, 3, 1, 2, so that two slices hold bins {0, 3} and {1, 2} and
// four slices one bin each.  Every aligned run of 64 or 128 slots still starts
// with its largest n_active (a bin's groups stay in descending order, and with
// one group per bin the pairs are (g0, g3) and (g1, g2)), which the kernels use
// to bound their rows.  Only for a batch without padding slots (B == Bp) of 4
// or a multiple of 8 groups; otherwise the sorted order stays.  Results per
// trajectory do not depend on its slot.
static std::vector<int> balanced_group_order(const std::vector<int>& sorted, int B, int Bp)
{
    const int G = B / 64;
    if (B != Bp || B % 64 != 0 || !(G == 4 || (G >= 8 && G % 8 == 0))) return sorted;
    std::vector<std::vector<int>> bin(4);
    for (int g = 0; g < G; ++g) {
        const int r = g % 8;
        bin[r < 4 ? r : 7 - r].push_back(g);
    }
    std::vector<int> out;
    out.reserve(B);
    for (int k : {0, 3, 1, 2})
        for (int g : bin[k])
            for (int q = 0; q < 64; ++q) out.push_back(sorted[g * 64 + q]);
    return out;
}

// basis vector groups the batched update keeps cached (dq::kUpdCacheBytes):
// a vector holds the active sites of every slot, so the budget is divided by
// their sum (the rows a slot does not use are never read)
static void set_upd_cached(dq_system* sys)
{
    size_t sites = 0;
    for (int32_t na : sys->nact_slot) sites += (size_t)na;
    const size_t group = 4 * std::max<size_t>(sites, 1) * (sys->d.mixed ? sizeof(float2) : sizeof(double2));
    sys->d.upd_cached = (int)std::min<size_t>(dq::kMaxM / 4, dq::kUpdCacheBytes / group);
}

static void release_row_blocks(dq_system* sys, RowBlocks* b)
{
    sys->release(b->ptr);
    sys->release(b->col);
    sys->release(b->val32);
    sys->release(b->val64);
    sys->release(b->lo16);
}

// bfloat16 bits of f, rounded to nearest even (f finite)
static uint16_t bf16_bits(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// the matrix-core SpMM reads x rows and dense block values through buffer
// resources with 32-bit byte offsets: one Krylov vector (n x padded batch,
// 8 bytes per entry) and the union slots (64 bytes each) must stay below 2^31
static bool mfma_offsets_fit(int64_t n, int64_t Bp, size_t union_slots)
{
    const uint64_t lim = (uint64_t)1 << 31;
    return (uint64_t)n * (uint64_t)Bp * 8 < lim && (uint64_t)union_slots * 64 < lim;
}

// R = 16 (the matrix-core SpMM, four union columns per step, four steps in
// flight): every block's union is padded to a multiple of 16 with zero-valued
// copies of its first column.
static dq_status make_row_blocks(dq_system* sys, int R, const std::vector<int32_t>& rowptr_v,
                                 const std::vector<int32_t>& col_v, const std::vector<double>& val_v, RowBlocks* out)
{
    const int n = sys->n, nblk = (n + R - 1) / R;
    const int32_t* rowptr = rowptr_v.data();
    const int32_t* col = col_v.data();
    const double* val = val_v.data();
    std::vector<int32_t> rbp(nblk + 1, 0), rbc;
    std::vector<double> rbv64;
    std::vector<int32_t> u;
    for (int k = 0; k < nblk; ++k) {
        const int r0 = k * R, r1 = std::min(n, r0 + R);
        u.clear();
        for (int i = r0; i < r1; ++i) u.insert(u.end(), col + rowptr[i], col + rowptr[i + 1]);
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        const size_t base = rbc.size();
        rbc.insert(rbc.end(), u.begin(), u.end());
        rbv64.resize(rbc.size() * R, 0.0);
        for (int i = r0; i < r1; ++i)
            for (int32_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
                const size_t pos = std::lower_bound(u.begin(), u.end(), col[p]) - u.begin();
                rbv64[(base + pos) * R + (i - r0)] = val[p];
            }
        if (R == 16) {
            while ((rbc.size() - base) % 16) rbc.push_back(u[0]);
            rbv64.resize(rbc.size() * R, 0.0);
        }
        rbp[k + 1] = (int32_t)rbc.size();
    }
    if (R == 16 && !mfma_offsets_fit(n, sys->d.Bp, rbc.size()))
        return fail(DQ_ERR_INVALID_ARG, "the matrix-core SpMM (spmm_rows_per_wave = 16) addresses a Krylov vector "
                                        "and the block values with 32-bit offsets: n * padded batch * 8 and "
                                        "64 * union slots must stay below 2^31 bytes");
    std::vector<float> rbv32(rbv64.size());
    for (size_t q = 0; q < rbv64.size(); ++q) rbv32[q] = (float)rbv64[q];
    // R = 16: the FP32 values are a fixed perturbation of the couplings (up to
    // half an FP32 ulp each), which the matrix-core SpMM takes back with one
    // v_mfma_f32_16x16x16bf16 per 16 union slots on the low parts
    // val - (double)val32, stored in that instruction's operand order: group G
    // of slots 16 G .. 16 G + 15, lane i + 16 kq, element q = row i, slot
    // 16 G + 4 q + kq (the slot lane group kq multiplies in K-step q)
    std::vector<uint16_t> lo16;
    if (R == 16) {
        const size_t ng = rbc.size() / 16;
        lo16.assign(std::max<size_t>(ng, 1) * 256, 0);
        for (size_t G = 0; G < ng; ++G)
            for (int l = 0; l < 64; ++l)
                for (int q = 0; q < 4; ++q) {
                    const size_t e = (16 * G + 4 * q + (l >> 4)) * 16 + (l & 15);
                    lo16[G * 256 + 4 * l + q] = bf16_bits((float)(rbv64[e] - (double)rbv32[e]));
                }
    }
    // the new blocks are complete before the old ones are released
    Pending pend(sys);
    RowBlocks nb;
    DQ_TRY(pend.alloc(&nb.ptr, rbp.size()));
    DQ_TRY(pend.alloc(&nb.col, std::max<size_t>(rbc.size(), 1)));
    DQ_TRY(pend.alloc(&nb.val32, std::max<size_t>(rbv32.size(), 1)));
    DQ_TRY(pend.alloc(&nb.val64, std::max<size_t>(rbv64.size(), 1)));
    if (R == 16) DQ_TRY(pend.alloc(&nb.lo16, lo16.size()));
    HIP_TRY(hipMemcpyAsync(nb.ptr, rbp.data(), sizeof(int32_t) * rbp.size(), hipMemcpyHostToDevice, sys->stream));
    if (!rbc.empty()) {
        HIP_TRY(hipMemcpyAsync(nb.col, rbc.data(), sizeof(int32_t) * rbc.size(), hipMemcpyHostToDevice, sys->stream));
        HIP_TRY(hipMemcpyAsync(nb.val32, rbv32.data(), sizeof(float) * rbv32.size(), hipMemcpyHostToDevice, sys->stream));
        HIP_TRY(hipMemcpyAsync(nb.val64, rbv64.data(), sizeof(double) * rbv64.size(), hipMemcpyHostToDevice, sys->stream));
        if (R == 16)
            HIP_TRY(hipMemcpyAsync(nb.lo16, lo16.data(), sizeof(uint16_t) * lo16.size(), hipMemcpyHostToDevice,
                                   sys->stream));
    }
    HIP_TRY(hipStreamSynchronize(sys->stream));
    pend.keep();
    release_row_blocks(sys, out);
    *out = nb;
    return DQ_OK;
}

// The tile and slice counts every SpMM consumer uses: the row-blocked
// kernels' tiles (4 row blocks of rb rows), and the tiles whose partial sums
// the SpMM leaves in part_s (the ring kernel's row ranges when it is on).
static void refresh_spmm_tiles(dq_system* sys)
{
    Dev& d = sys->d;
    if (!d.plain) {
        d.rt_s = 4 * d.rb;
        d.ntile_s = (sys->n + d.rt_s - 1) / d.rt_s;
    }
    d.ntile_p = d.ring_rows > 0 ? (sys->n + d.ring_rows - 1) / d.ring_rows : d.ntile_s;
}

// row blocks of R rows for the Hamiltonian (from the host CSR kept by
// set_hamiltonian) and, with Taylor-Krylov on, for the auxiliary matrix
static dq_status build_row_blocks(dq_system* sys, int R)
{
    Dev& d = sys->d;
    RowBlocks a, t;
    DQ_TRY(make_row_blocks(sys, R, sys->rowptr, sys->col, sys->val, &a));
    if (sys->tk.p > 0) {
        dq_status st = make_row_blocks(sys, R, sys->tk.rowptr_h, sys->tk.col_h, sys->tk.val_h, &t);
        if (st != DQ_OK) {
            release_row_blocks(sys, &a);
            return st;
        }
        release_row_blocks(sys, &sys->tk.rb);
        sys->tk.rb = t;
    }
    RowBlocks old;
    old.ptr = const_cast<int32_t*>(d.rb_ptr);
    old.col = const_cast<int32_t*>(d.rb_col);
    old.val32 = const_cast<float*>(d.rb_val32);
    old.val64 = const_cast<double*>(d.rb_val64);
    old.lo16 = const_cast<uint16_t*>(d.rb_lo16);
    release_row_blocks(sys, &old);
    d.rb_ptr = a.ptr; d.rb_col = a.col; d.rb_val32 = a.val32; d.rb_val64 = a.val64; d.rb_lo16 = a.lo16;
    d.rb = R;
    refresh_spmm_tiles(sys);
    return DQ_OK;
}

static void release_ring(dq_system* sys, RingLists* r)
{
    sys->release(r->item);
    sys->release(r->a);
    sys->release(r->alo);
    sys->release(r->off);
    sys->release(r->win);
    *r = RingLists{};
}

// Lists of the ring SpMM for the Hamiltonian's host CSR.  Fails (nothing
// changed) when the band is too wide: the x rows that one 128-row chunk and
// the next reference must fit the ring of kRingCap rows.
static dq_status make_ring_lists(dq_system* sys, RingLists* out)
{
    constexpr int CAP = dq::kRingCap, CH = dq::kRingChunk;
    const int n = sys->n, nblk = (n + 3) / 4, nchunk = (n + CH - 1) / CH;
    const int32_t* rowptr = sys->rowptr.data();
    const int32_t* col = sys->col.data();
    const double* val = sys->val.data();
    // windows of the 128-row chunks: their own rows and every column they
    // reference, made monotone (hi a running maximum, lo a running minimum
    // from the end) and rounded to groups of 4 rows
    std::vector<int2> win(nchunk);
    for (int c = 0; c < nchunk; ++c) {
        const int r0 = c * CH, r1 = std::min(n, r0 + CH);
        int lo = r0, hi = r1;
        for (int i = r0; i < r1; ++i)
            for (int32_t p = rowptr[i]; p < rowptr[i + 1]; ++p) { lo = std::min(lo, col[p]); hi = std::max(hi, col[p] + 1); }
        win[c] = make_int2(lo, hi);
    }
    const int nr4 = (n + 3) & ~3;
    for (int c = 1; c < nchunk; ++c) win[c].y = std::max(win[c].y, win[c - 1].y);
    for (int c = nchunk - 2; c >= 0; --c) win[c].x = std::min(win[c].x, win[c + 1].x);
    for (int c = 0; c < nchunk; ++c) {
        win[c].x &= ~3;
        win[c].y = std::min(nr4, (win[c].y + 3) & ~3);
        const int span = (c + 1 < nchunk ? win[c + 1].y : win[c].y) - win[c].x;
        if (span > CAP)
            return fail(DQ_ERR_INVALID_ARG, "ring SpMM: rows " + std::to_string(c * CH) + ".. reference x rows " +
                                                std::to_string(span) + " apart, more than its LDS ring of " +
                                                std::to_string(CAP) + " rows (the Hamiltonian's band is too wide)");
    }
    // per 4-row block: the pairs (2p, 2p + 1) covering its column union
    // (a block without couplings takes the pair of its own first rows, so
    // that its item still forms w = d x), padded to a multiple of 4 with
    // zero-coupling copies of the last pair (rows of its window), in items of
    // up to 16 pairs
    std::vector<int32_t> item(nblk + 1, 0);
    std::vector<float> a;
    std::vector<uint16_t> alo;
    std::vector<uint32_t> off;
    std::vector<int32_t> pr;
    std::vector<std::pair<int32_t, double>> row[4];
    int nit = 0;
    for (int k = 0; k < nblk; ++k) {
        const int r0 = 4 * k;
        pr.clear();
        for (int i = 0; i < 4; ++i) {
            row[i].clear();
            if (r0 + i >= n) continue;
            for (int32_t p = rowptr[r0 + i]; p < rowptr[r0 + i + 1]; ++p) {
                row[i].push_back({col[p], val[p]});
                pr.push_back(col[p] >> 1);
            }
            std::sort(row[i].begin(), row[i].end());
        }
        std::sort(pr.begin(), pr.end());
        pr.erase(std::unique(pr.begin(), pr.end()), pr.end());
        if (pr.empty()) pr.push_back(r0 >> 1);
        const int np = (int)pr.size(), np4 = (np + 3) & ~3, nk = (np4 + 15) / 16;
        item[k] = nit;
        a.resize((size_t)(nit + nk) * 128, 0.f);
        alo.resize((size_t)(nit + nk) * 128, 0);
        off.resize((size_t)(nit + nk) * 16, 0);
        for (int q = 0; q < 16 * nk; ++q) {
            const int it = nit + q / 16, sl = q % 16;
            const int pair = q < np ? pr[q] : pr[np - 1];
            off[(size_t)it * 16 + sl] = (uint32_t)(((2 * pair) & (CAP - 1)) * 4);
            if (q >= np) continue;
            for (int h = 0; h < 2; ++h)
                for (int i = 0; i < 4; ++i) {
                    const auto e = std::lower_bound(row[i].begin(), row[i].end(), std::make_pair(2 * pair + h, -1e308));
                    const double v = (e != row[i].end() && e->first == 2 * pair + h) ? e->second : 0.0;
                    const float v32 = (float)v;
                    a[((size_t)it * 64 + 32 * h + 4 * (sl % 8) + i) * 2 + sl / 8] = v32;
                    alo[(size_t)it * 128 + (((sl / 4) * 2 + h) * 4 + i) * 4 + sl % 4] = bf16_bits((float)(v - (double)v32));
                }
        }
        // the item's pair count and the block's last item in the low bits of
        // its first offset (a multiple of 8)
        for (int t = 0; t < nk; ++t) {
            const int cnt = std::min(16, np4 - 16 * t);
            off[(size_t)(nit + t) * 16] |= (uint32_t)(cnt / 4 - 1) | (t == nk - 1 ? 4u : 0u);
        }
        nit += nk;
    }
    item[nblk] = nit;
    if ((uint64_t)nit * 512 >= ((uint64_t)1 << 31))
        return fail(DQ_ERR_INVALID_ARG, "ring SpMM: the coupling items exceed 2^31 bytes");
    Pending pend(sys);
    RingLists r;
    DQ_TRY(pend.alloc(&r.item, (size_t)nblk + 1));
    DQ_TRY(pend.alloc(&r.a, a.size()));
    DQ_TRY(pend.alloc(&r.alo, alo.size() / 4));
    DQ_TRY(pend.alloc(&r.off, off.size() / 4));
    DQ_TRY(pend.alloc(&r.win, (size_t)nchunk));
    HIP_TRY(hipMemcpyAsync(r.item, item.data(), sizeof(int32_t) * item.size(), hipMemcpyHostToDevice, sys->stream));
    HIP_TRY(hipMemcpyAsync(r.a, a.data(), sizeof(float) * a.size(), hipMemcpyHostToDevice, sys->stream));
    HIP_TRY(hipMemcpyAsync(r.alo, alo.data(), sizeof(uint16_t) * alo.size(), hipMemcpyHostToDevice, sys->stream));
    HIP_TRY(hipMemcpyAsync(r.off, off.data(), sizeof(uint32_t) * off.size(), hipMemcpyHostToDevice, sys->stream));
    HIP_TRY(hipMemcpyAsync(r.win, win.data(), sizeof(int2) * win.size(), hipMemcpyHostToDevice, sys->stream));
    HIP_TRY(hipStreamSynchronize(sys->stream));
    if (!sys->ring_lds_set) {
        const size_t need = dq::spmm_ring_lds_bytes(16, dq::kRingMaxRows), lim = ring_lds_limit();
        if (need > lim)
            return fail(DQ_ERR_INVALID_ARG, "ring SpMM: needs " + std::to_string(need) + " bytes of LDS per "
                                                "workgroup, the device allows " + std::to_string(lim));
        const hipError_t e = dq::spmm_ring_set_lds_limit();
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(DQ_ERR_INVALID_ARG, std::string("ring SpMM: the LDS limit cannot be raised: ") +
                                                hipGetErrorString(e));
        }
        sys->ring_lds_set = true;
    }
    pend.keep();
    release_ring(sys, out);
    *out = r;
    return DQ_OK;
}

// Point the device view at the ring lists (ring_rows > 0 and lists built) or
// switch the ring SpMM off.
static void attach_ring(dq_system* sys)
{
    Dev& d = sys->d;
    const RingLists& r = sys->ring;
    d.ring_rows = (sys->ring_rows > 0 && r.item) ? sys->ring_rows : 0;
    d.ring_w = sys->ring_waves;
    d.ring_item = r.item;
    d.ring_a = r.a;
    d.ring_alo = r.alo;
    d.ring_off = r.off;
    d.ring_win = r.win;
    refresh_spmm_tiles(sys);
}

extern "C" {

dq_status dq_system_create(const dq_config* cfg, dq_system** out)
{
    if (!cfg || !out) return fail(DQ_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (cfg->n_sites <= 0 || cfg->n_traj <= 0) return fail(DQ_ERR_INVALID_ARG, "n_sites and n_traj must be > 0");
    if (cfg->krylov_m < 2 || cfg->krylov_m > dq::kMaxM) return fail(DQ_ERR_INVALID_ARG, "krylov_m must be in [2, 32]");
    if (cfg->krylov_substeps < 1) return fail(DQ_ERR_INVALID_ARG, "krylov_substeps must be >= 1");
    if (cfg->ou_substeps < 1) return fail(DQ_ERR_INVALID_ARG, "ou_substeps must be >= 1");
    if (cfg->precision != DQ_PREC_FP64 && cfg->precision != DQ_PREC_MIXED)
        return fail(DQ_ERR_INVALID_ARG, "unknown precision");
    if (cfg->noise_model < DQ_NOISE_OU || cfg->noise_model > DQ_NOISE_NONE)
        return fail(DQ_ERR_INVALID_ARG, "unknown noise model");
    if (!(cfg->dt_fs > 0.0)) return fail(DQ_ERR_INVALID_ARG, "dt_fs must be > 0");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(DQ_ERR_NO_DEVICE, "no HIP device");
    dq_system* s = new dq_system();
    s->cfg = *cfg;
    s->n = cfg->n_sites;
    s->B = cfg->n_traj;
    s->m = cfg->krylov_m;
    s->K = cfg->krylov_substeps;
    const bool plain = (s->B == 1 && !cfg->force_batched && cfg->precision == DQ_PREC_FP64);
    s->Bp = plain ? 1 : ((s->B + 63) / 64) * 64;
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete s; return fail(DQ_ERR_HIP, std::string("stream: ") + hipGetErrorString(e)); }
    hipEventCreate(&s->ev0);
    hipEventCreate(&s->ev1);
    e = hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming);
    for (int k = 0; k < dq_system::kMaxSlices - 1 && e == hipSuccess; ++k) {
        e = hipStreamCreateWithFlags(&s->side[k], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev_join[k], hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        delete s;
        return fail(DQ_ERR_HIP, std::string("stream/event: ") + hipGetErrorString(e));
    }
    Dev& d = s->d;
    d.n = s->n; d.Bp = s->Bp; d.B = s->B; d.m = s->m;
    d.mixed = cfg->precision == DQ_PREC_MIXED;
    d.plain = plain;
    d.rt_e = plain ? 1024 : 128;
    d.rb = 4;
    d.aug = 2;
    d.spmm_tpl = (d.mixed && d.Bp % 128 == 0) ? 2 : 1;
    d.spmm_phases = 1;
    d.spmm_pipe = 0;
    d.s0 = 0;
    d.sn = d.Bp;
    d.rt_s = plain ? 64 : 4 * d.rb;      // batched: one row block per wave
    d.ntile_e = (d.n + d.rt_e - 1) / d.rt_e;
    d.rt_o = plain ? d.rt_e : dq::kRowsMin;
    d.ntile_o = (d.n + d.rt_o - 1) / d.rt_o;
    d.ntile_s = (d.n + d.rt_s - 1) / d.rt_s;
    d.ntile_p = d.ntile_s;
    d.ring_w = s->ring_waves;
    d.S = cfg->ou_substeps;
    d.noise_model = cfg->noise_model;
    d.dt = cfg->dt_fs;
    d.seed = cfg->seed;
    // systemClass.cu:993-994: subdt = dt / S, sqrt(subdt)
    d.subdt = cfg->dt_fs / (double)cfg->ou_substeps;
    d.sqrt_subdt = std::sqrt(d.subdt);
    d.g = cfg->coupling_scale;
    const double hbar = cfg->hbar_cm_fs > 0.0 ? cfg->hbar_cm_fs : reference_hbar();
    s->hbar = hbar;
    s->K0 = cfg->krylov_substeps;
    // propagate.cu:90-92: expCoeff = -i (1/hbar) dt, here per Krylov substep
    d.theta = -((-1. / hbar) * (cfg->dt_fs / (double)cfg->krylov_substeps));
    d.tol = cfg->breakdown_tol > 0.0 ? cfg->breakdown_tol : (d.mixed ? 1e-5 : 1e-10);
    d.sigma_static = cfg->sigma_static;
    const size_t nB = (size_t)d.n * d.Bp;
    dq_status st = DQ_OK;
    double2* psi = nullptr;
    void* V = nullptr;
    void* diag = nullptr;
#define ALLOC(p, cnt) do { if ((st = s->alloc(p, cnt)) != DQ_OK) { delete s; return st; } } while (0)
    ALLOC(&psi, nB);
    d.vs = (nB + 63) / 64 * 64 + 32;
    if (d.mixed) { float2* v; ALLOC(&v, d.vs * (s->m + 1)); V = v; float* dg; ALLOC(&dg, nB); diag = dg; }
    else { double2* v; ALLOC(&v, d.vs * (s->m + 1)); V = v; double* dg; ALLOC(&dg, nB); diag = dg; }
    d.psi = psi; d.V = V; d.diag = diag;
    ALLOC(&d.de, nB);
    ALLOC(&d.e0, nB);
    ALLOC(&d.sc, (size_t)(s->m + 1) * d.Bp);
    ALLOC(&d.alpha, (size_t)s->m * d.Bp);
    ALLOC(&d.beta, (size_t)s->m * d.Bp);
    ALLOC(&d.hco, (size_t)dq::kMaxM * dq::kMaxM * d.Bp);
    ALLOC(&d.gram, (size_t)(s->m + 1) * d.Bp);
    ALLOC(&d.nrm2, (size_t)d.Bp);
    ALLOC(&d.ycoef, (size_t)s->m * d.Bp);
    ALLOC(&d.err_sum, (size_t)d.Bp);
    ALLOC(&d.broke, (size_t)d.Bp);
    ALLOC(&d.status, 1);
    size_t nq = std::max<size_t>(2 * (s->m + 1) + 1, dq::kMaxGroups + 3);
    // partial sums sized for the smallest tiles the tuning allows
    const size_t te_max = plain ? (size_t)d.ntile_e : (size_t)((d.n + dq::kRowsMin - 1) / dq::kRowsMin);
    const size_t ts_max = plain ? (size_t)d.ntile_s : (size_t)((d.n + 4 * dq::kRBMin - 1) / (4 * dq::kRBMin));
    size_t part = std::max(nq * te_max, (size_t)2 * ts_max) * d.Bp;
    ALLOC(&d.part, part);
    ALLOC(&d.part_s, (size_t)2 * ts_max * d.Bp);
    ALLOC(&d.red, (nq + 2) * d.Bp);
    ALLOC(&d.step, 1);
    // trajectories: ids 0..B-1, all sites active; padding trajectories inactive
    {
        int32_t* na; uint64_t* ti;
        ALLOC(&na, (size_t)d.Bp);
        ALLOC(&ti, (size_t)d.Bp);
        std::vector<int32_t> hn(d.Bp, 0);
        std::vector<uint64_t> ht(d.Bp, 0);
        for (int b = 0; b < s->B; ++b) { hn[b] = s->n; ht[b] = (uint64_t)b; }
        s->slot.resize(s->B);
        for (int b = 0; b < s->B; ++b) s->slot[b] = b;
        s->nact_slot = hn;
        set_upd_cached(s);
        hipMemcpyAsync(na, hn.data(), sizeof(int32_t) * d.Bp, hipMemcpyHostToDevice, s->stream);
        hipMemcpyAsync(ti, ht.data(), sizeof(uint64_t) * d.Bp, hipMemcpyHostToDevice, s->stream);
        d.nact = na; d.traj_id = ti;
    }
#undef ALLOC
    dq::launch_set_unit_norm(d, s->stream);
    e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) { delete s; return fail(DQ_ERR_HIP, std::string("create: ") + hipGetErrorString(e)); }
    *out = s;
    return DQ_OK;
}

dq_status dq_system_destroy(dq_system* sys)
{
    if (!sys) return DQ_OK;
    hipStreamSynchronize(sys->stream);
    delete sys;
    return DQ_OK;
}

dq_status dq_sync(dq_system* sys)
{
    if (!sys) return fail(DQ_ERR_INVALID_ARG, "null system");
    HIP_TRY(hipStreamSynchronize(sys->stream));
    int flags = 0;
    HIP_TRY(hipMemcpy(&flags, sys->d.status, sizeof(int), hipMemcpyDeviceToHost));
    if (flags & dq::kStatusScaling) return fail(DQ_ERR_SCALING_OVERFLOW, "Pade scaling exponent s >= 50");
    if (flags & dq::kStatusNonfinite) return fail(DQ_ERR_NONFINITE, "non-finite error estimate");
    return DQ_OK;
}

dq_status dq_system_set_hamiltonian_csr(dq_system* sys, int32_t n, const int32_t* rowptr,
                                        const int32_t* col, const double* val, const double* e_site)
{
    if (!sys || !rowptr) return fail(DQ_ERR_INVALID_ARG, "null argument");
    if (n != sys->n) return fail(DQ_ERR_INVALID_ARG, "n does not match n_sites");
    if (rowptr[0] != 0) return fail(DQ_ERR_INVALID_ARG, "rowptr[0] must be 0");
    const int64_t nnz = rowptr[n];
    for (int i = 0; i < n; ++i)
        if (rowptr[i + 1] < rowptr[i]) return fail(DQ_ERR_INVALID_ARG, "rowptr not monotone");
    if (nnz > 0 && (!col || !val)) return fail(DQ_ERR_INVALID_ARG, "null col/val");
    for (int64_t p = 0; p < nnz; ++p) {
        if (col[p] < 0 || col[p] >= n) return fail(DQ_ERR_INVALID_ARG, "column index out of range");
        if (!std::isfinite(val[p])) return fail(DQ_ERR_NONFINITE, "non-finite coupling");
    }
    for (int i = 0; i < n; ++i)
        for (int32_t p = rowptr[i]; p < rowptr[i + 1]; ++p)
            if (col[p] == i) return fail(DQ_ERR_INVALID_ARG, "diagonal entries belong in e_site");
    Dev& d = sys->d;
    sys->drop_graph();
    HIP_TRY(hipStreamSynchronize(sys->stream));
    // the new CSR, row blocks and ring lists are complete before anything
    // old is released: a failure leaves the previous Hamiltonian in place
    Pending pend(sys);
    int32_t *rp, *cl;
    double *v64, *es;
    float* v32;
    DQ_TRY(pend.alloc(&rp, (size_t)n + 1));
    DQ_TRY(pend.alloc(&cl, (size_t)std::max<int64_t>(nnz, 1)));
    DQ_TRY(pend.alloc(&v64, (size_t)std::max<int64_t>(nnz, 1)));
    DQ_TRY(pend.alloc(&v32, (size_t)std::max<int64_t>(nnz, 1)));
    DQ_TRY(pend.alloc(&es, (size_t)n));
    std::vector<float> f32(nnz);
    for (int64_t p = 0; p < nnz; ++p) f32[p] = (float)val[p];
    std::vector<double> esv(n, 0.0);
    if (e_site) for (int i = 0; i < n; ++i) esv[i] = e_site[i];
    HIP_TRY(hipMemcpyAsync(rp, rowptr, sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, sys->stream));
    if (nnz > 0) {
        HIP_TRY(hipMemcpyAsync(cl, col, sizeof(int32_t) * nnz, hipMemcpyHostToDevice, sys->stream));
        HIP_TRY(hipMemcpyAsync(v64, val, sizeof(double) * nnz, hipMemcpyHostToDevice, sys->stream));
        HIP_TRY(hipMemcpyAsync(v32, f32.data(), sizeof(float) * nnz, hipMemcpyHostToDevice, sys->stream));
    }
    HIP_TRY(hipMemcpyAsync(es, esv.data(), sizeof(double) * n, hipMemcpyHostToDevice, sys->stream));
    HIP_TRY(hipStreamSynchronize(sys->stream));
    // host copies swapped in for the builders, swapped back on failure
    std::vector<int32_t> rp_h(rowptr, rowptr + n + 1), col_h(col, col + nnz);
    std::vector<double> val_h(val, val + nnz);
    std::swap(sys->rowptr, rp_h);
    std::swap(sys->col, col_h);
    std::swap(sys->val, val_h);
    auto restore_host = [&]() {
        std::swap(sys->rowptr, rp_h);
        std::swap(sys->col, col_h);
        std::swap(sys->val, val_h);
    };
    RingLists nr;
    if (sys->ring_rows > 0) {
        dq_status st = make_ring_lists(sys, &nr);
        if (st != DQ_OK) { restore_host(); return st; }
    }
    dq_status st = build_row_blocks(sys, d.rb);
    if (st != DQ_OK) {
        release_ring(sys, &nr);
        restore_host();
        return st;
    }
    pend.keep();
    if (sys->ring_rows > 0) {
        release_ring(sys, &sys->ring);
        sys->ring = nr;
        attach_ring(sys);
    }
    sys->release(*(int32_t**)&d.rowptr);
    sys->release(*(int32_t**)&d.col);
    sys->release(*(double**)&d.val64);
    sys->release(*(float**)&d.val32);
    sys->release(*(double**)&d.e_site);
    d.rowptr = rp; d.col = cl; d.val64 = v64; d.val32 = v32; d.e_site = es;
    sys->have_h = true;
    return DQ_OK;
}

static void release_tk(dq_system* sys)
{
    Dev& d = sys->d;
    sys->release(sys->tk.rowptr);
    sys->release(sys->tk.col);
    sys->release(sys->tk.val64);
    sys->release(sys->tk.val32);
    release_row_blocks(sys, &sys->tk.rb);
    sys->release(sys->tk.diag);
    sys->release(d.tk_t);
    sys->release(d.tk_u);
    sys->release(*(double2**)&d.tk_coef);
    sys->release(d.tk_nrm);
    sys->release(d.tk_err);
    sys->tk.p = sys->tk.m = 0;
    d.tk_p = 0;
}

// taylorKrylovClass.cu:133-215 (set-up), the tkmatfile / tk_p / tk_m_aux keys
// (systemClass.cu:455-469)
dq_status dq_system_set_taylor_krylov(dq_system* sys, int32_t p, int32_t m_aux, const int32_t* rowptr,
                                      const int32_t* col, const double* val, const double* diag)
{
    if (!sys) return fail(DQ_ERR_INVALID_ARG, "null system");
    Dev& d = sys->d;
    const int n = sys->n;
    if (p != 0 && sys->ktol > 0.0) return fail(DQ_ERR_INVALID_ARG, "substep control is for the Lanczos propagator");
    if (p != 0) {
        if (p < 2 || m_aux < 2 || m_aux > sys->m || m_aux + p > dq::kMaxM + 2)
            return fail(DQ_ERR_INVALID_ARG, "Taylor-Krylov needs p >= 2, 2 <= m_aux <= m and m_aux + p <= 34");
        if (!rowptr || !col || !val || !diag) return fail(DQ_ERR_INVALID_ARG, "null auxiliary matrix");
        if (rowptr[0] != 0) return fail(DQ_ERR_INVALID_ARG, "rowptr[0] != 0");
        for (int i = 0; i < n; ++i) {
            if (rowptr[i + 1] < rowptr[i]) return fail(DQ_ERR_INVALID_ARG, "rowptr not monotone");
            if (!std::isfinite(diag[i])) return fail(DQ_ERR_NONFINITE, "non-finite auxiliary diagonal");
            for (int32_t q = rowptr[i]; q < rowptr[i + 1]; ++q) {
                if (col[q] < 0 || col[q] >= n) return fail(DQ_ERR_INVALID_ARG, "column out of range");
                if (col[q] == i) return fail(DQ_ERR_INVALID_ARG, "diagonal entries belong in diag");
                if (!std::isfinite(val[q])) return fail(DQ_ERR_NONFINITE, "non-finite auxiliary coupling");
            }
        }
    }
    sys->drop_graph();
    hipStreamSynchronize(sys->stream);
    release_tk(sys);
    if (p == 0) return DQ_OK;
    const int64_t nnz = rowptr[n];
    const size_t nB = (size_t)n * d.Bp;
    sys->tk.rowptr_h.assign(rowptr, rowptr + n + 1);
    sys->tk.col_h.assign(col, col + nnz);
    sys->tk.val_h.assign(val, val + nnz);
    DQ_TRY(sys->alloc(&sys->tk.rowptr, (size_t)n + 1));
    DQ_TRY(sys->alloc(&sys->tk.col, (size_t)std::max<int64_t>(nnz, 1)));
    DQ_TRY(sys->alloc(&sys->tk.val64, (size_t)std::max<int64_t>(nnz, 1)));
    DQ_TRY(sys->alloc(&sys->tk.val32, (size_t)std::max<int64_t>(nnz, 1)));
    std::vector<float> f32(nnz);
    for (int64_t q = 0; q < nnz; ++q) f32[q] = (float)val[q];
    HIP_TRY(hipMemcpyAsync(sys->tk.rowptr, rowptr, sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, sys->stream));
    if (nnz > 0) {
        HIP_TRY(hipMemcpyAsync(sys->tk.col, col, sizeof(int32_t) * nnz, hipMemcpyHostToDevice, sys->stream));
        HIP_TRY(hipMemcpyAsync(sys->tk.val64, val, sizeof(double) * nnz, hipMemcpyHostToDevice, sys->stream));
        HIP_TRY(hipMemcpyAsync(sys->tk.val32, f32.data(), sizeof(float) * nnz, hipMemcpyHostToDevice, sys->stream));
    }
    // the auxiliary diagonal for every trajectory slot, in the basis precision
    if (d.mixed) {
        float* dg;
        DQ_TRY(sys->alloc(&dg, nB));
        std::vector<float> h(nB);
        for (int i = 0; i < n; ++i)
            for (int b = 0; b < d.Bp; ++b) h[(size_t)i * d.Bp + b] = (float)diag[i];
        HIP_TRY(hipMemcpyAsync(dg, h.data(), sizeof(float) * nB, hipMemcpyHostToDevice, sys->stream));
        HIP_TRY(hipStreamSynchronize(sys->stream));
        sys->tk.diag = dg;
    } else {
        double* dg;
        DQ_TRY(sys->alloc(&dg, nB));
        std::vector<double> h(nB);
        for (int i = 0; i < n; ++i)
            for (int b = 0; b < d.Bp; ++b) h[(size_t)i * d.Bp + b] = diag[i];
        HIP_TRY(hipMemcpyAsync(dg, h.data(), sizeof(double) * nB, hipMemcpyHostToDevice, sys->stream));
        HIP_TRY(hipStreamSynchronize(sys->stream));
        sys->tk.diag = dg;
    }
    // c^k / k! of c = -i tau / hbar by repeated complex products from 1
    // (taylorKrylovClass.cu:200-214, cppUtility.cpp:52-79); k! in double,
    // where the reference's int factorial would overflow past 12!
    std::vector<double2> coef(p);
    {
        const double cr = 0.0, ci = -d.theta;
        double zr = 1.0, zi = 0.0, fact = 1.0;
        for (int k = 0; k < p; ++k) {
            if (k > 0) {
                const double r = zr * cr - zi * ci, im = zr * ci + zi * cr;
                zr = r;
                zi = im;
                fact *= (double)k;
            }
            coef[k] = make_double2(zr / fact, zi / fact);
        }
    }
    double2* dc;
    DQ_TRY(sys->alloc(&dc, (size_t)p));
    HIP_TRY(hipMemcpyAsync(dc, coef.data(), sizeof(double2) * p, hipMemcpyHostToDevice, sys->stream));
    DQ_TRY(sys->alloc(&d.tk_t, nB));
    DQ_TRY(sys->alloc(&d.tk_u, 2 * nB));
    DQ_TRY(sys->alloc(&d.tk_nrm, (size_t)d.Bp));
    DQ_TRY(sys->alloc(&d.tk_err, (size_t)d.Bp));
    d.tk_coef = dc;
    sys->tk.p = p;
    sys->tk.m = m_aux;
    DQ_TRY(make_row_blocks(sys, d.rb, sys->tk.rowptr_h, sys->tk.col_h, sys->tk.val_h, &sys->tk.rb));
    d.tk_p = p;
    HIP_TRY(hipStreamSynchronize(sys->stream));
    return DQ_OK;
}

dq_status dq_system_set_ou_params(dq_system* sys, const double* tau_fs, const double* sigma)
{
    if (!sys || !tau_fs || !sigma) return fail(DQ_ERR_INVALID_ARG, "null argument");
    const int n = sys->n;
    std::vector<double> it(n), co(n), sg(n);
    for (int i = 0; i < n; ++i) {
        if (!(tau_fs[i] > 0.0) || !(sigma[i] >= 0.0)) return fail(DQ_ERR_INVALID_ARG, "tau must be > 0, sigma >= 0");
        // systemClass.cu:1033-1034
        it[i] = 1. / tau_fs[i];
        co[i] = sigma[i] * std::sqrt(2. / tau_fs[i]);
        sg[i] = sigma[i];
        if (sigma[i] == 0.0) { it[i] = 0.0; co[i] = 0.0; }
    }
    Dev& d = sys->d;
    sys->drop_graph();
    HIP_TRY(hipStreamSynchronize(sys->stream));
    Pending pend(sys);
    double *a, *b, *c;
    DQ_TRY(pend.alloc(&a, n));
    DQ_TRY(pend.alloc(&b, n));
    DQ_TRY(pend.alloc(&c, n));
    HIP_TRY(hipMemcpyAsync(a, it.data(), sizeof(double) * n, hipMemcpyHostToDevice, sys->stream));
    HIP_TRY(hipMemcpyAsync(b, co.data(), sizeof(double) * n, hipMemcpyHostToDevice, sys->stream));
    HIP_TRY(hipMemcpyAsync(c, sg.data(), sizeof(double) * n, hipMemcpyHostToDevice, sys->stream));
    HIP_TRY(hipStreamSynchronize(sys->stream));
    pend.keep();
    sys->release(*(double**)&d.inv_tau);
    sys->release(*(double**)&d.coeff);
    sys->release(*(double**)&d.ou_sigma);
    d.inv_tau = a; d.coeff = b; d.ou_sigma = c;
    sys->have_ou = true;
    return DQ_OK;
}

dq_status dq_system_set_hsr_params(dq_system* sys, const double* sigma)
{
    if (!sys || !sigma) return fail(DQ_ERR_INVALID_ARG, "null argument");
    const int n = sys->n;
    for (int i = 0; i < n; ++i)
        if (!(sigma[i] >= 0.0) || !std::isfinite(sigma[i])) return fail(DQ_ERR_INVALID_ARG, "sigma must be finite and >= 0");
    Dev& d = sys->d;
    sys->drop_graph();
    HIP_TRY(hipStreamSynchronize(sys->stream));
    Pending pend(sys);
    double* a;
    DQ_TRY(pend.alloc(&a, n));
    HIP_TRY(hipMemcpyAsync(a, sigma, sizeof(double) * n, hipMemcpyHostToDevice, sys->stream));
    HIP_TRY(hipStreamSynchronize(sys->stream));
    pend.keep();
    sys->release(*(double**)&d.hsr_sigma);
    d.hsr_sigma = a;
    sys->have_hsr = true;
    return DQ_OK;
}

dq_status dq_system_set_zz_params(dq_system* sys, int32_t nosc, const double* omega_cm, const double* coeff)
{
    if (!sys || !omega_cm || !coeff) return fail(DQ_ERR_INVALID_ARG, "null argument");
    if (nosc < 1) return fail(DQ_ERR_INVALID_ARG, "nosc must be >= 1");
    // includeFiles.cu:16 (INVFS_TO_INVCM): omega in rad/fs = omega_cm / 5308.84
    const double conv = 5308.84, pi = 3.14159265358979323846;
    std::vector<double> w(nosc);
    for (int k = 0; k < nosc; ++k) {
        if (!std::isfinite(omega_cm[k]) || !std::isfinite(coeff[k]) || omega_cm[k] < 0.0)
            return fail(DQ_ERR_INVALID_ARG, "omega must be finite and >= 0, coeff finite");
        w[k] = omega_cm[k] / conv / (2 * pi);
    }
    Dev& d = sys->d;
    sys->drop_graph();
    HIP_TRY(hipStreamSynchronize(sys->stream));
    Pending pend(sys);
    double *a, *c;
    DQ_TRY(pend.alloc(&a, nosc));
    DQ_TRY(pend.alloc(&c, nosc));
    HIP_TRY(hipMemcpyAsync(a, w.data(), sizeof(double) * nosc, hipMemcpyHostToDevice, sys->stream));
    HIP_TRY(hipMemcpyAsync(c, coeff, sizeof(double) * nosc, hipMemcpyHostToDevice, sys->stream));
    HIP_TRY(hipStreamSynchronize(sys->stream));
    pend.keep();
    sys->release(*(double**)&d.zz_w2pi);
    sys->release(*(double**)&d.zz_coeff);
    d.zz_w2pi = a;
    d.zz_coeff = c;
    d.zz_nosc = nosc;
    sys->have_zz = true;
    return DQ_OK;
}

dq_status dq_zz_oscillators(int32_t nosc, const double* omega_cm, const double* J, double domega,
                            double temperature_K, double* coeff)
{
    if (!omega_cm || !J || !coeff) return fail(DQ_ERR_INVALID_ARG, "null argument");
    if (nosc < 1 || !(domega > 0.0) || !(temperature_K > 0.0))
        return fail(DQ_ERR_INVALID_ARG, "nosc >= 1, domega > 0 and temperature > 0 required");
    // systemClass.cu:1118-1145 (hbar = 1 in cm^-1 units), includeFiles.cu:13
    const double kB = 0.695035, pi = 3.14159265358979323846;
    for (int k = 0; k < nosc; ++k) {
        const double G = J[k] / (1 - std::exp(-1. / (kB * temperature_K) * omega_cm[k]));
        coeff[k] = std::sqrt(2.) * std::sqrt(G * domega / pi);
        if (!std::isfinite(coeff[k])) return fail(DQ_ERR_NONFINITE, "non-finite amplitude (omega = 0 or J < 0?)");
    }
    return DQ_OK;
}

dq_status dq_system_set_trajectories(dq_system* sys, const uint64_t* traj_ids, const int32_t* n_active)
{
    if (!sys) return fail(DQ_ERR_INVALID_ARG, "null system");
    Dev& d = sys->d;
    const int B = sys->B;
    std::vector<int32_t> na(B);
    for (int b = 0; b < B; ++b) {
        na[b] = n_active ? n_active[b] : sys->n;
        if (na[b] < 1 || na[b] > sys->n) return fail(DQ_ERR_INVALID_ARG, "n_active out of range");
    }
    // slots ordered by n_active, largest first (stable, so equal sizes keep
    // the caller's order); the plain path has one slot
    std::vector<int> order(B);
    for (int b = 0; b < B; ++b) order[b] = b;
    if (!d.plain) std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return na[x] > na[y]; });
    if (!d.plain) order = balanced_group_order(order, B, d.Bp);
    std::vector<int32_t> hn(d.Bp, 0);
    std::vector<uint64_t> ht(d.Bp, 0);
    for (int k = 0; k < B; ++k) {
        const int b = order[k];
        sys->slot[b] = k;
        ht[k] = traj_ids ? traj_ids[b] : (uint64_t)b;
        hn[k] = na[b];
    }
    sys->nact_slot = hn;
    set_upd_cached(sys);
    sys->drop_graph();
    HIP_TRY(hipMemcpyAsync((void*)d.nact, hn.data(), sizeof(int32_t) * d.Bp, hipMemcpyHostToDevice, sys->stream));
    HIP_TRY(hipMemcpyAsync((void*)d.traj_id, ht.data(), sizeof(uint64_t) * d.Bp, hipMemcpyHostToDevice, sys->stream));
    // rows a slot does not use must read 0 in every vector; the kernels never
    // write rows that are inactive for a whole slice, so clear them all here
    const size_t nB = (size_t)d.n * d.Bp;
    const size_t vb = d.mixed ? sizeof(float2) : sizeof(double2);
    HIP_TRY(hipMemsetAsync(d.V, 0, vb * d.vs * (sys->m + 1), sys->stream));
    HIP_TRY(hipMemsetAsync(d.psi, 0, sizeof(double2) * nB, sys->stream));
    if (d.tk_t) {
        HIP_TRY(hipMemsetAsync(d.tk_t, 0, sizeof(double2) * nB, sys->stream));
        HIP_TRY(hipMemsetAsync(d.tk_u, 0, sizeof(double2) * nB * 2, sys->stream));
    }
    HIP_TRY(hipStreamSynchronize(sys->stream));
    sys->have_disorder = false;
    sys->have_state = false;
    return DQ_OK;
}

dq_status dq_system_reset_disorder(dq_system* sys)
{
    if (!sys) return fail(DQ_ERR_INVALID_ARG, "null system");
    if (!sys->have_h || !sys->noise_ready()) return fail(DQ_ERR_STATE, "set the Hamiltonian and noise parameters first");
    Dev& d = sys->d;
    HIP_TRY(dq::launch_init_disorder(d, sys->stream));
    HIP_TRY(hipMemsetAsync(d.err_sum, 0, sizeof(double) * d.Bp, sys->stream));
    HIP_TRY(hipMemsetAsync(d.step, 0, sizeof(int64_t), sys->stream));
    HIP_TRY(hipMemsetAsync(d.status, 0, sizeof(int), sys->stream));
    HIP_TRY(hipStreamSynchronize(sys->stream));
    sys->step = 0;
    sys->have_disorder = true;
    return DQ_OK;
}

// The reference normalises every initial state on the host
// (systemClass.cu:1383-1481, normalizeHostCol).  Here |psi_b|^2 of each
// trajectory's active sites is handed to the device as the norm the first
// prologue divides by, the same renormalisation every later step applies.
dq_status dq_system_set_state(dq_system* sys, const double* psi)
{
    if (!sys || !psi) return fail(DQ_ERR_INVALID_ARG, "null argument");
    Dev& d = sys->d;
    const int n = sys->n, Bp = d.Bp;
    std::vector<double> h((size_t)2 * n * Bp, 0.0), n2(Bp, 1.0);
    for (int b = 0; b < sys->B; ++b) {
        const int sb = sys->slot[b], na = sys->nact_slot[sb];
        double s2 = 0.0;
        for (int i = 0; i < n; ++i) {
            const double re = psi[((size_t)b * n + i) * 2], im = psi[((size_t)b * n + i) * 2 + 1];
            if (!std::isfinite(re) || !std::isfinite(im)) return fail(DQ_ERR_NONFINITE, "non-finite state");
            if (i >= na) continue;                 // sites beyond n_active are not part of b
            h[((size_t)i * Bp + sb) * 2] = re;
            h[((size_t)i * Bp + sb) * 2 + 1] = im;
            s2 += re * re + im * im;
        }
        if (!(s2 > 0.0)) return fail(DQ_ERR_INVALID_ARG, "trajectory " + std::to_string(b) + ": zero state on its active sites");
        n2[sb] = s2;
    }
    HIP_TRY(hipMemcpyAsync(d.psi, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, sys->stream));
    HIP_TRY(hipMemcpyAsync(d.nrm2, n2.data(), sizeof(double) * Bp, hipMemcpyHostToDevice, sys->stream));
    HIP_TRY(hipStreamSynchronize(sys->stream));
    sys->have_state = true;
    return DQ_OK;
}

dq_status dq_system_get_state(dq_system* sys, double* psi)
{
    if (!sys || !psi) return fail(DQ_ERR_INVALID_ARG, "null argument");
    Dev& d = sys->d;
    const int n = sys->n, Bp = d.Bp;
    std::vector<double> h((size_t)2 * n * Bp);
    HIP_TRY(hipStreamSynchronize(sys->stream));
    HIP_TRY(hipMemcpy(h.data(), d.psi, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    for (int b = 0; b < sys->B; ++b) {
        const int sb = sys->slot[b];
        for (int i = 0; i < n; ++i) {
            psi[((size_t)b * n + i) * 2] = h[((size_t)i * Bp + sb) * 2];
            psi[((size_t)b * n + i) * 2 + 1] = h[((size_t)i * Bp + sb) * 2 + 1];
        }
    }
    return DQ_OK;
}

// Resume point of a run (no counterpart in the reference, whose runs restart
// from the initial state): the state psi as the next step will use it, the
// OU offsets, the error sums and the step counter.  Together with the
// deterministic static disorder (dq_system_reset_disorder) and the Philox
// counters of the noise, they fix every later step bit for bit.
dq_status dq_system_snapshot(dq_system* sys, double* psi, double* noise, double* err_sum, int64_t* step)
{
    if (!sys || !psi || !noise || !step) return fail(DQ_ERR_INVALID_ARG, "null argument");
    if (!sys->have_state) return fail(DQ_ERR_STATE, "no state to snapshot (call dq_system_set_state)");
    Dev& d = sys->d;
    const int n = sys->n, Bp = d.Bp;
    // a normalisation still pending from dq_system_set_state is applied now,
    // exactly as the next step's prologue would (afterwards that is a no-op)
    HIP_TRY(dq::launch_prologue(d, 1, 0, 0, 0, sys->stream));
    HIP_TRY(dq::launch_set_unit_norm(d, sys->stream));
    DQ_TRY(dq_system_get_state(sys, psi));
    std::vector<double> h((size_t)n * Bp);
    HIP_TRY(hipMemcpy(h.data(), d.de, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    for (int b = 0; b < sys->B; ++b)
        for (int i = 0; i < n; ++i) noise[(size_t)b * n + i] = h[(size_t)i * Bp + sys->slot[b]];
    if (err_sum) DQ_TRY(dq_get_error_sum(sys, err_sum));
    int64_t s = 0;
    HIP_TRY(hipMemcpy(&s, d.step, sizeof(int64_t), hipMemcpyDeviceToHost));
    *step = s;
    return DQ_OK;
}

dq_status dq_system_restore(dq_system* sys, const double* psi, const double* noise, const double* err_sum,
                            int64_t step)
{
    if (!sys || !psi || !noise) return fail(DQ_ERR_INVALID_ARG, "null argument");
    if (!sys->have_h || !sys->noise_ready() || !sys->have_disorder)
        return fail(DQ_ERR_STATE, "set the Hamiltonian and noise parameters and call dq_system_reset_disorder first");
    if (step < 0) return fail(DQ_ERR_INVALID_ARG, "step < 0");
    Dev& d = sys->d;
    const int n = sys->n, Bp = d.Bp;
    const size_t nB = (size_t)n * Bp;
    std::vector<double> hp(2 * nB, 0.0), hd(nB, 0.0), e0(nB), es(Bp, 0.0), n2(Bp, 1.0);
    for (int b = 0; b < sys->B; ++b) {
        const int sb = sys->slot[b], na = sys->nact_slot[sb];
        for (int i = 0; i < na; ++i) {
            const double re = psi[((size_t)b * n + i) * 2], im = psi[((size_t)b * n + i) * 2 + 1];
            const double de = noise[(size_t)b * n + i];
            if (!std::isfinite(re) || !std::isfinite(im) || !std::isfinite(de))
                return fail(DQ_ERR_NONFINITE, "non-finite snapshot");
            hp[((size_t)i * Bp + sb) * 2] = re;
            hp[((size_t)i * Bp + sb) * 2 + 1] = im;
            hd[(size_t)i * Bp + sb] = de;
        }
        if (err_sum) es[sb] = err_sum[b];
    }
    HIP_TRY(hipStreamSynchronize(sys->stream));
    HIP_TRY(hipMemcpy(e0.data(), d.e0, sizeof(double) * nB, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpyAsync(d.psi, hp.data(), sizeof(double) * hp.size(), hipMemcpyHostToDevice, sys->stream));
    HIP_TRY(hipMemcpyAsync(d.nrm2, n2.data(), sizeof(double) * Bp, hipMemcpyHostToDevice, sys->stream));
    HIP_TRY(hipMemcpyAsync(d.de, hd.data(), sizeof(double) * nB, hipMemcpyHostToDevice, sys->stream));
    HIP_TRY(hipMemcpyAsync(d.err_sum, es.data(), sizeof(double) * Bp, hipMemcpyHostToDevice, sys->stream));
    HIP_TRY(hipMemcpyAsync(d.step, &step, sizeof(int64_t), hipMemcpyHostToDevice, sys->stream));
    // the diagonal the last noise update wrote: static energy + offset, in the
    // basis precision (active sites; the rest as reset_disorder left them)
    if (d.mixed) {
        std::vector<float> dg(nB);
        HIP_TRY(hipMemcpy(dg.data(), d.diag, sizeof(float) * nB, hipMemcpyDeviceToHost));
        for (int b = 0; b < sys->B; ++b) {
            const int sb = sys->slot[b], na = sys->nact_slot[sb];
            for (int i = 0; i < na; ++i) dg[(size_t)i * Bp + sb] = (float)(hd[(size_t)i * Bp + sb] + e0[(size_t)i * Bp + sb]);
        }
        HIP_TRY(hipMemcpyAsync(d.diag, dg.data(), sizeof(float) * nB, hipMemcpyHostToDevice, sys->stream));
        HIP_TRY(hipStreamSynchronize(sys->stream));
    } else {
        std::vector<double> dg(nB);
        HIP_TRY(hipMemcpy(dg.data(), d.diag, sizeof(double) * nB, hipMemcpyDeviceToHost));
        for (int b = 0; b < sys->B; ++b) {
            const int sb = sys->slot[b], na = sys->nact_slot[sb];
            for (int i = 0; i < na; ++i) dg[(size_t)i * Bp + sb] = hd[(size_t)i * Bp + sb] + e0[(size_t)i * Bp + sb];
        }
        HIP_TRY(hipMemcpyAsync(d.diag, dg.data(), sizeof(double) * nB, hipMemcpyHostToDevice, sys->stream));
        HIP_TRY(hipStreamSynchronize(sys->stream));
    }
    sys->step = step;
    sys->have_state = true;
    return DQ_OK;
}

dq_status dq_system_set_observables(dq_system* sys, const int32_t* group, int32_t ngroups, int32_t every,
                                    int64_t max_steps, const double* bra)
{
    if (!sys) return fail(DQ_ERR_INVALID_ARG, "null system");
    Dev& d = sys->d;
    const int n = sys->n;
    if (max_steps < 0 || every < 0) return fail(DQ_ERR_INVALID_ARG, "negative every/max_steps");
    if (every > 0 && (!group || ngroups <= 0)) return fail(DQ_ERR_INVALID_ARG, "groups required");
    int site_mode = 0;
    if (every > 0) {
        site_mode = (ngroups == n);
        for (int i = 0; i < n && site_mode; ++i) if (group[i] != i) site_mode = 0;
        if (!site_mode && ngroups > dq::kMaxGroups)
            return fail(DQ_ERR_INVALID_ARG, "at most 8 population groups (or one per site)");
        for (int i = 0; i < n; ++i)
            if (group[i] >= ngroups) return fail(DQ_ERR_INVALID_ARG, "group index out of range");
    }
    sys->drop_graph();
    HIP_TRY(hipStreamSynchronize(sys->stream));
    Pending pend(sys);
    const int64_t nrec = every > 0 ? max_steps / every + 1 : 0;
    int32_t* gp = nullptr;
    double* pops = nullptr;
    double2 *br = nullptr, *corr = nullptr;
    if (every > 0) {
        DQ_TRY(pend.alloc(&gp, n));
        DQ_TRY(pend.alloc(&pops, (size_t)nrec * (site_mode ? n : ngroups) * d.Bp));
        HIP_TRY(hipMemcpyAsync(gp, group, sizeof(int32_t) * n, hipMemcpyHostToDevice, sys->stream));
    }
    if (bra) {
        DQ_TRY(pend.alloc(&br, n));
        DQ_TRY(pend.alloc(&corr, (size_t)(max_steps + 1) * d.Bp));
        HIP_TRY(hipMemcpyAsync(br, bra, sizeof(double2) * n, hipMemcpyHostToDevice, sys->stream));
    }
    HIP_TRY(hipStreamSynchronize(sys->stream));
    pend.keep();
    sys->release(*(int32_t**)&d.group);
    sys->release(d.pops);
    sys->release(*(double2**)&d.bra);
    sys->release(*(int32_t**)&d.bra_idx);
    sys->release(d.corr);
    sys->release(d.partic);
    d.group = gp;
    d.pops = pops;
    d.bra = br;
    d.corr = corr;
    d.ngroups = ngroups;
    d.every = every;
    d.site_mode = site_mode;
    d.max_steps = max_steps;
    sys->ngroups = ngroups;
    sys->every = every;
    sys->site_mode = site_mode;
    sys->max_steps = max_steps;
    sys->obs_cap_rec = nrec;
    sys->have_bra = bra != nullptr;
    return DQ_OK;
}

dq_status dq_system_set_participation(dq_system* sys, int32_t on)
{
    if (!sys) return fail(DQ_ERR_INVALID_ARG, "null system");
    Dev& d = sys->d;
    sys->drop_graph();
    hipStreamSynchronize(sys->stream);
    sys->release(d.partic);
    if (on) DQ_TRY(sys->alloc(&d.partic, (size_t)(sys->max_steps + 1) * d.Bp));
    return DQ_OK;
}

// ---------------------------------------------------------------------------
// step sequence
// ---------------------------------------------------------------------------
}  // extern "C"

// brackets one launch with events when profiling
template <typename F>
static hipError_t timed(dq_system* sys, int cls, F&& launch)
{
    if (!sys->profiling) return launch();
    while (sys->ev_pool.size() < sys->ev_used + 2) {
        hipEvent_t ev;
        hipError_t e = hipEventCreate(&ev);
        if (e != hipSuccess) return e;
        sys->ev_pool.push_back(ev);
    }
    hipEvent_t a = sys->ev_pool[sys->ev_used], b = sys->ev_pool[sys->ev_used + 1];
    sys->ev_used += 2;
    sys->ev_class.push_back(cls);
    if (cls == DQ_K_UPDATE) {
        // the update launchers (batched and plain) take the events into the
        // launch itself: the dominant kernel's own duration, as a kernel
        // tracer reports it, without the dispatch gap of events around it
        dq::g_kev = {a, b, false};
        hipError_t e = launch();
        const bool used = dq::g_kev.used;
        dq::g_kev = {};
        if (e == hipSuccess && !used) e = hipErrorInvalidHandle;   // a launcher that ignores them
        return e;
    }
    hipError_t e = hipEventRecord(a, sys->stream);
    if (e == hipSuccess) e = launch();
    if (e == hipSuccess) e = hipEventRecord(b, sys->stream);
    return e;
}

static hipError_t collect_timings(dq_system* sys)
{
    if (!sys->profiling) return hipSuccess;
    hipError_t e = hipStreamSynchronize(sys->stream);
    if (e != hipSuccess) return e;
    for (size_t k = 0; k < sys->ev_class.size(); ++k) {
        float ms = 0.f;
        e = hipEventElapsedTime(&ms, sys->ev_pool[2 * k], sys->ev_pool[2 * k + 1]);
        if (e != hipSuccess) return e;
        sys->kms[sys->ev_class[k]] += ms;
        sys->kcount[sys->ev_class[k]] += 1;
    }
    sys->ev_class.clear();
    sys->ev_used = 0;
    return hipSuccess;
}

extern "C" {

// Lanczos decomposition of d's matrix on basis slot 0, then the Pade kernel
static hipError_t enqueue_lanczos_pade(dq_system* sys, const Dev& d, hipStream_t st)
{
    hipError_t e;
    // batched: the Gram partials of update j-1 are reduced by leading blocks of
    // the matrix-core SpMM of iteration j and the scalars kernel sums the SpMM
    // tiles itself, or both are reduced by one launch after the other SpMMs
    // (the last update's |v_m|^2 is summed by the Pade kernel); plain: the
    // scalars kernel sums the SpMM tiles and each update is followed by its
    // reduction
    for (int j = 0; j < d.m; ++j) {
        if ((e = timed(sys, DQ_K_SPMM, [&] { return dq::launch_spmm(d, j, 0, st); })) != hipSuccess) return e;
        if (!d.plain && !dq::spmm_fuses_gram(d) &&
            (e = timed(sys, DQ_K_REDUCE, [&] { return dq::launch_reduce_all(d, j, st); })) != hipSuccess) return e;
        if ((e = timed(sys, DQ_K_SCALARS, [&] { return dq::launch_scalars_h(d, j, st); })) != hipSuccess) return e;
        if ((e = timed(sys, DQ_K_UPDATE, [&] { return dq::launch_update(d, j, st); })) != hipSuccess) return e;
        if (d.plain && (e = timed(sys, DQ_K_REDUCE, [&] { return dq::launch_reduce(d, 2 * (j + 1) + 1, d.ntile_e, st); })) != hipSuccess) return e;
    }
    if ((e = timed(sys, DQ_K_SPMM_LAST, [&] { return dq::launch_spmm(d, d.m, 1, st); })) != hipSuccess) return e;
    return timed(sys, DQ_K_PADE, [&] { return dq::launch_pade(d, st); });
}

// the auxiliary Krylov part of a Taylor-Krylov cycle: d with the auxiliary
// matrix, its m and augmentation p
static Dev tk_dev(const dq_system* sys, const Dev& d)
{
    Dev a = d;
    a.m = sys->tk.m;
    a.aug = sys->tk.p;
    a.tk = 1;
    a.rowptr = sys->tk.rowptr;
    a.col = sys->tk.col;
    a.val64 = sys->tk.val64;
    a.val32 = sys->tk.val32;
    a.rb_ptr = sys->tk.rb.ptr;
    a.rb_col = sys->tk.rb.col;
    a.rb_val32 = sys->tk.rb.val32;
    a.rb_val64 = sys->tk.rb.val64;
    a.rb_lo16 = sys->tk.rb.lo16;
    a.diag = sys->tk.diag;
    a.g = 1.0;
    // the auxiliary matrix goes through the row-blocked SpMM
    a.ring_rows = 0;
    a.ntile_p = a.ntile_s;
    return a;
}

static hipError_t enqueue_cycle_dev(dq_system* sys, const Dev& d, bool first_in_step, bool last_in_step,
                                    bool obs, hipStream_t st, bool advance)
{
    hipError_t e;
    const bool tk = sys->tk.p > 0;
    if ((e = timed(sys, DQ_K_PROLOGUE, [&] { return dq::launch_prologue(d, 1, first_in_step, first_in_step && obs, tk ? 0 : 1, st); })) != hipSuccess) return e;
    if (first_in_step && obs && (e = timed(sys, DQ_K_NORM, [&] { return dq::launch_reduce_obs(d, st); })) != hipSuccess) return e;
    if (!tk) {
        if ((e = enqueue_lanczos_pade(sys, d, st)) != hipSuccess) return e;
        if ((e = timed(sys, DQ_K_COMBINE, [&] { return dq::launch_combine(d, st); })) != hipSuccess) return e;
        return timed(sys, DQ_K_NORM, [&] { return dq::launch_finish_norm(d, (last_in_step && advance) ? 1 : 0, st); });
    }
    // Taylor-Krylov (taylorKrylovClass.cu:217-434): Taylor part with the
    // Hamiltonian, Krylov part with the auxiliary matrix
    for (int k = 1; k <= sys->tk.p; ++k)
        if ((e = timed(sys, DQ_K_SPMM, [&] { return dq::launch_tk_power(d, k, st); })) != hipSuccess) return e;
    if ((e = timed(sys, DQ_K_NORM, [&] { return dq::launch_tk_start(d, st); })) != hipSuccess) return e;
    const Dev a = tk_dev(sys, d);
    if ((e = enqueue_lanczos_pade(sys, a, st)) != hipSuccess) return e;
    if ((e = timed(sys, DQ_K_COMBINE, [&] { return dq::launch_tk_combine(a, st); })) != hipSuccess) return e;
    return timed(sys, DQ_K_NORM, [&] { return dq::launch_finish_norm(a, (last_in_step && advance) ? 1 : 0, st); });
}

static hipError_t enqueue_cycle_dev(dq_system* sys, const Dev& d, bool first_in_step, bool last_in_step,
                                    bool obs)
{
    return enqueue_cycle_dev(sys, d, first_in_step, last_in_step, obs, sys->stream, true);
}

static hipError_t enqueue_cycle(dq_system* sys, bool first_in_step, bool last_in_step)
{
    const Dev& d = sys->d;
    return enqueue_cycle_dev(sys, d, first_in_step, last_in_step, dq::obs_on(d));
}

static hipError_t enqueue_step(dq_system* sys)
{
    hipError_t e = hipSuccess;
    for (int k = 0; k < sys->K; ++k)
        if ((e = enqueue_cycle(sys, k == 0, k == sys->K - 1)) != hipSuccess) return e;
    return e;
}

// One step with the trajectory slots split into `slices` equal slices, each on
// its own stream (fork/join by events; used inside graph capture, where the
// chains of small latency-bound kernels fill each other's gaps).  The
// per-trajectory arithmetic is unchanged, so results are bitwise those of one
// stream.  The step counter moves once, after all slices.
static hipError_t enqueue_step_slices(dq_system* sys)
{
    const Dev& d = sys->d;
    const bool obs = dq::obs_on(d);
    const int ns = sys->slices;
    hipError_t e;
    if ((e = hipEventRecord(sys->ev_fork, sys->stream)) != hipSuccess) return e;
    for (int h = 1; h < ns; ++h)
        if ((e = hipStreamWaitEvent(sys->side[h - 1], sys->ev_fork, 0)) != hipSuccess) return e;
    for (int h = 0; h < ns; ++h) {
        Dev dh = d;
        dh.sn = d.Bp / ns;
        dh.s0 = h * dh.sn;
        hipStream_t st = h ? sys->side[h - 1] : sys->stream;
        for (int k = 0; k < sys->K; ++k)
            if ((e = enqueue_cycle_dev(sys, dh, k == 0, k == sys->K - 1, obs, st, false)) != hipSuccess) return e;
    }
    for (int h = 1; h < ns; ++h) {
        if ((e = hipEventRecord(sys->ev_join[h - 1], sys->side[h - 1])) != hipSuccess) return e;
        if ((e = hipStreamWaitEvent(sys->stream, sys->ev_join[h - 1], 0)) != hipSuccess) return e;
    }
    return dq::launch_advance_step(d, sys->stream);
}

static hipError_t enqueue_finalize(dq_system* sys)
{
    const Dev& d = sys->d;
    const bool obs = dq::obs_on(d);
    hipError_t e = dq::launch_prologue(d, 1, 0, obs, 0, sys->stream);
    if (e != hipSuccess) return e;
    if (obs && (e = dq::launch_reduce_obs(d, sys->stream)) != hipSuccess) return e;
    return dq::launch_set_unit_norm(d, sys->stream);
}

static dq_status check_ready(dq_system* sys)
{
    if (!sys) return fail(DQ_ERR_INVALID_ARG, "null system");
    if (!sys->have_h || !sys->noise_ready()) return fail(DQ_ERR_STATE, "Hamiltonian/noise parameters not set");
    if (!sys->have_disorder) return fail(DQ_ERR_STATE, "call dq_system_reset_disorder first");
    if (!sys->have_state) return fail(DQ_ERR_STATE, "call dq_system_set_state first");
    return DQ_OK;
}

dq_status dq_update_noise(dq_system* sys)
{
    DQ_TRY(check_ready(sys));
    HIP_TRY(dq::launch_prologue(sys->d, 0, 1, 0, 0, sys->stream));
    HIP_TRY(dq::launch_advance_step(sys->d, sys->stream));
    sys->step += 1;
    return dq_sync(sys);
}

dq_status dq_propagate_step(dq_system* sys)
{
    return dq_run(sys, 1, 0);
}

}  // extern "C"

// nsteps steps enqueued (the graph captured first when asked for), timed by
// events; no finalisation
static dq_status run_steps(dq_system* sys, int64_t nsteps, int use_graph, float* ms_out)
{
    hipStream_t st = sys->stream;
    if (use_graph && nsteps > 0 && !sys->graph) {
        hipGraph_t g;
        HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipError_t e = sys->slices > 1 ? enqueue_step_slices(sys) : enqueue_step(sys);
        hipError_t e2 = hipStreamEndCapture(st, &g);
        if (e != hipSuccess || e2 != hipSuccess)
            return fail(DQ_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
        e = hipGraphInstantiate(&sys->graph, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e != hipSuccess) { sys->graph = nullptr; return fail(DQ_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e)); }
    }
    HIP_TRY(hipEventRecord(sys->ev0, st));
    for (int64_t k = 0; k < nsteps; ++k) {
        if (use_graph) HIP_TRY(hipGraphLaunch(sys->graph, st));
        else HIP_TRY(enqueue_step(sys));
        if (sys->profiling && sys->ev_used > 4096) HIP_TRY(collect_timings(sys));
    }
    HIP_TRY(hipEventRecord(sys->ev1, st));
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(collect_timings(sys));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, sys->ev0, sys->ev1));
    *ms_out = ms;
    return DQ_OK;
}

// K Krylov cycles per dt from now on: tau = dt / K
static void set_substeps(dq_system* sys, int K)
{
    if (K == sys->K) return;
    sys->drop_graph();
    sys->K = K;
    // propagate.cu:90-92, per Krylov substep, the expression of dq_system_create
    sys->d.theta = -((-1. / sys->hbar) * (sys->cfg.dt_fs / (double)K));
}

// the state a chunk starts from, for a redo with more substeps
static dq_status chunk_save(dq_system* sys, bool restore)
{
    Dev& d = sys->d;
    hipStream_t st = sys->stream;
    const size_t nB = (size_t)sys->n * d.Bp, vb = d.mixed ? sizeof(float) : sizeof(double);
    auto cp = [&](void* dst, const void* src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st);
    };
    auto& c = sys->ck;
    if (!restore) {
        HIP_TRY(cp(c.psi, d.psi, sizeof(double2) * nB));
        HIP_TRY(cp(c.de, d.de, sizeof(double) * nB));
        HIP_TRY(cp(c.diag, d.diag, vb * nB));
        HIP_TRY(cp(c.err, d.err_sum, sizeof(double) * d.Bp));
        HIP_TRY(cp(c.nrm2, d.nrm2, sizeof(double) * d.Bp));
        HIP_TRY(cp(c.status, d.status, sizeof(int)));
        HIP_TRY(cp(c.step, d.step, sizeof(int64_t)));
    } else {
        HIP_TRY(cp(d.psi, c.psi, sizeof(double2) * nB));
        HIP_TRY(cp(d.de, c.de, sizeof(double) * nB));
        HIP_TRY(cp(d.diag, c.diag, vb * nB));
        HIP_TRY(cp(d.err_sum, c.err, sizeof(double) * d.Bp));
        HIP_TRY(cp(d.nrm2, c.nrm2, sizeof(double) * d.Bp));
        HIP_TRY(cp(d.status, c.status, sizeof(int)));
        HIP_TRY(cp(d.step, c.step, sizeof(int64_t)));
    }
    return DQ_OK;
}

// Substep control from the Expokit estimate (padeClass.cu:749-782,
// systemClass.cu:2055-2085): the reference sums the estimate; here a chunk
// of steps is accepted when K times the largest clamped estimate of any
// cycle of any trajectory stays within the tolerance per step, and redone
// from its start with twice the substeps otherwise (up to kmax).
static dq_status run_controlled(dq_system* sys, int64_t nsteps, int use_graph, float* ms_total)
{
    Dev& d = sys->d;
    hipStream_t st = sys->stream;
    std::vector<double> em(d.Bp);
    int64_t done = 0;
    float total = 0.f;
    while (done < nsteps) {
        const int64_t c = std::min<int64_t>(sys->kchunk, nsteps - done);
        DQ_TRY(chunk_save(sys, false));
        HIP_TRY(hipMemsetAsync(d.err_max, 0, sizeof(double) * d.Bp, st));
        float ms = 0.f;
        DQ_TRY(run_steps(sys, c, use_graph, &ms));
        total += ms;
        HIP_TRY(hipMemcpy(em.data(), d.err_max, sizeof(double) * d.Bp, hipMemcpyDeviceToHost));
        double emax = 0.0;
        for (int b = 0; b < sys->B; ++b) emax = std::max(emax, em[sys->slot[b]]);
        if (!(emax * sys->K <= sys->ktol) && 2 * sys->K <= sys->kmax) {
            DQ_TRY(chunk_save(sys, true));
            set_substeps(sys, 2 * sys->K);
            ++sys->k_rejects;
            continue;
        }
        done += c;
        sys->step += c;
    }
    sys->step -= nsteps;               // dq_run adds the whole run
    *ms_total = total;
    return DQ_OK;
}

extern "C" {

dq_status dq_run(dq_system* sys, int64_t nsteps, int32_t use_graph)
{
    DQ_TRY(check_ready(sys));
    if (nsteps < 0) return fail(DQ_ERR_INVALID_ARG, "nsteps < 0");
    if (dq::obs_on(sys->d) && sys->step + nsteps > sys->max_steps)
        return fail(DQ_ERR_INVALID_ARG, "run exceeds max_steps of the observables");
    if (sys->profiling) use_graph = 0;
    float ms = 0.f;
    if (sys->ktol > 0.0) DQ_TRY(run_controlled(sys, nsteps, use_graph, &ms));
    else DQ_TRY(run_steps(sys, nsteps, use_graph, &ms));
    HIP_TRY(enqueue_finalize(sys));
    HIP_TRY(hipStreamSynchronize(sys->stream));
    sys->last_ms = ms;
    sys->step += nsteps;
    return dq_sync(sys);
}

dq_status dq_system_set_krylov_control(dq_system* sys, double tol, int32_t k_max, int32_t chunk_steps)
{
    if (!sys) return fail(DQ_ERR_INVALID_ARG, "null system");
    Dev& d = sys->d;
    if (!(tol > 0.0)) {
        sys->ktol = 0.0;
        d.err_max = nullptr;
        sys->drop_graph();
        set_substeps(sys, sys->K0);
        return DQ_OK;
    }
    if (sys->tk.p > 0) return fail(DQ_ERR_INVALID_ARG, "substep control is for the Lanczos propagator");
    if (k_max < sys->K0 || chunk_steps < 1) return fail(DQ_ERR_INVALID_ARG, "need k_max >= krylov_substeps and chunk_steps >= 1");
    sys->drop_graph();
    HIP_TRY(hipStreamSynchronize(sys->stream));
    auto& c = sys->ck;
    if (!c.psi) {
        const size_t nB = (size_t)sys->n * d.Bp;
        Pending pend(sys);
        double* em;
        DQ_TRY(pend.alloc(&em, (size_t)d.Bp));
        DQ_TRY(pend.alloc(&c.psi, nB));
        DQ_TRY(pend.alloc(&c.de, nB));
        if (d.mixed) { float* f; DQ_TRY(pend.alloc(&f, nB)); c.diag = f; }
        else { double* f; DQ_TRY(pend.alloc(&f, nB)); c.diag = f; }
        DQ_TRY(pend.alloc(&c.err, (size_t)d.Bp));
        DQ_TRY(pend.alloc(&c.nrm2, (size_t)d.Bp));
        DQ_TRY(pend.alloc(&c.status, 1));
        DQ_TRY(pend.alloc(&c.step, 1));
        pend.keep();
        d.err_max = em;
        sys->ck_err_max = em;
    } else {
        d.err_max = sys->ck_err_max;
    }
    sys->ktol = tol;
    sys->kmax = k_max;
    sys->kchunk = chunk_steps;
    return DQ_OK;
}

dq_status dq_get_krylov_control(dq_system* sys, int32_t* k_now, int64_t* rejected_chunks)
{
    if (!sys || !k_now || !rejected_chunks) return fail(DQ_ERR_INVALID_ARG, "null argument");
    *k_now = sys->K;
    *rejected_chunks = sys->k_rejects;
    return DQ_OK;
}

dq_status dq_set_profiling(dq_system* sys, int32_t on)
{
    if (!sys) return fail(DQ_ERR_INVALID_ARG, "null system");
    sys->profiling = on != 0;
    for (int c = 0; c < DQ_K_CLASSES; ++c) { sys->kms[c] = 0.0; sys->kcount[c] = 0; }
    sys->ev_class.clear();
    sys->ev_used = 0;
    return DQ_OK;
}

dq_status dq_get_kernel_times(dq_system* sys, double* ms, int64_t* launches)
{
    if (!sys || !ms || !launches) return fail(DQ_ERR_INVALID_ARG, "null argument");
    for (int c = 0; c < DQ_K_CLASSES; ++c) { ms[c] = sys->kms[c]; launches[c] = sys->kcount[c]; }
    return DQ_OK;
}

dq_status dq_get_step(dq_system* sys, int64_t* step)
{
    if (!sys || !step) return fail(DQ_ERR_INVALID_ARG, "null argument");
    int64_t s = 0;
    HIP_TRY(hipStreamSynchronize(sys->stream));
    HIP_TRY(hipMemcpy(&s, sys->d.step, sizeof(int64_t), hipMemcpyDeviceToHost));
    *step = s;
    return DQ_OK;
}

dq_status dq_last_run_ms(dq_system* sys, double* ms)
{
    if (!sys || !ms) return fail(DQ_ERR_INVALID_ARG, "null argument");
    *ms = sys->last_ms;
    return DQ_OK;
}

dq_status dq_get_populations(dq_system* sys, double* pops, int64_t* nrec_out)
{
    if (!sys) return fail(DQ_ERR_INVALID_ARG, "null system");
    if (sys->every <= 0) return fail(DQ_ERR_STATE, "no population observables set");
    const Dev& d = sys->d;
    int64_t nrec = std::min<int64_t>(sys->step / sys->every + 1, sys->obs_cap_rec);
    if (nrec_out) *nrec_out = nrec;
    if (!pops) return DQ_OK;
    const int G = sys->site_mode ? sys->n : sys->ngroups;
    std::vector<double> h((size_t)nrec * G * d.Bp);
    HIP_TRY(hipStreamSynchronize(sys->stream));
    HIP_TRY(hipMemcpy(h.data(), d.pops, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    for (int b = 0; b < sys->B; ++b) {
        const int sb = sys->slot[b];
        for (int64_t r = 0; r < nrec; ++r)
            for (int g = 0; g < G; ++g)
                pops[((size_t)b * nrec + r) * G + g] = h[((size_t)r * G + g) * d.Bp + sb];
    }
    return DQ_OK;
}

dq_status dq_get_corr(dq_system* sys, double* corr)
{
    if (!sys || !corr) return fail(DQ_ERR_INVALID_ARG, "null argument");
    if (!sys->have_bra) return fail(DQ_ERR_STATE, "no bra set");
    const Dev& d = sys->d;
    const int64_t T = std::min(sys->step, sys->max_steps);
    std::vector<double2> h((size_t)(T + 1) * d.Bp);
    HIP_TRY(hipStreamSynchronize(sys->stream));
    HIP_TRY(hipMemcpy(h.data(), d.corr, sizeof(double2) * h.size(), hipMemcpyDeviceToHost));
    for (int b = 0; b < sys->B; ++b) {
        const int sb = sys->slot[b];
        for (int64_t k = 0; k <= T; ++k) {
            corr[((size_t)b * (T + 1) + k) * 2] = h[(size_t)k * d.Bp + sb].x;
            corr[((size_t)b * (T + 1) + k) * 2 + 1] = h[(size_t)k * d.Bp + sb].y;
        }
    }
    return DQ_OK;
}

dq_status dq_get_participation(dq_system* sys, double* pr)
{
    if (!sys || !pr) return fail(DQ_ERR_INVALID_ARG, "null argument");
    const Dev& d = sys->d;
    if (!d.partic) return fail(DQ_ERR_STATE, "participation ratio not enabled");
    const int64_t T = std::min(sys->step, sys->max_steps);
    std::vector<double> h((size_t)(T + 1) * d.Bp);
    HIP_TRY(hipStreamSynchronize(sys->stream));
    HIP_TRY(hipMemcpy(h.data(), d.partic, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    for (int b = 0; b < sys->B; ++b) {
        const int sb = sys->slot[b];
        for (int64_t k = 0; k <= T; ++k) pr[(size_t)b * (T + 1) + k] = h[(size_t)k * d.Bp + sb];
    }
    return DQ_OK;
}

dq_status dq_get_error_sum(dq_system* sys, double* err)
{
    if (!sys || !err) return fail(DQ_ERR_INVALID_ARG, "null argument");
    std::vector<double> h(sys->d.Bp);
    HIP_TRY(hipStreamSynchronize(sys->stream));
    HIP_TRY(hipMemcpy(h.data(), sys->d.err_sum, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    for (int b = 0; b < sys->B; ++b) err[b] = h[sys->slot[b]];
    return DQ_OK;
}

dq_status dq_get_diagonal(dq_system* sys, double* diag)
{
    if (!sys || !diag) return fail(DQ_ERR_INVALID_ARG, "null argument");
    const Dev& d = sys->d;
    const size_t nB = (size_t)sys->n * d.Bp;
    HIP_TRY(hipStreamSynchronize(sys->stream));
    if (d.mixed) {
        std::vector<float> h(nB);
        HIP_TRY(hipMemcpy(h.data(), d.diag, sizeof(float) * nB, hipMemcpyDeviceToHost));
        for (int b = 0; b < sys->B; ++b)
            for (int i = 0; i < sys->n; ++i) diag[(size_t)b * sys->n + i] = h[(size_t)i * d.Bp + sys->slot[b]];
    } else {
        std::vector<double> h(nB);
        HIP_TRY(hipMemcpy(h.data(), d.diag, sizeof(double) * nB, hipMemcpyDeviceToHost));
        for (int b = 0; b < sys->B; ++b)
            for (int i = 0; i < sys->n; ++i) diag[(size_t)b * sys->n + i] = h[(size_t)i * d.Bp + sys->slot[b]];
    }
    return DQ_OK;
}

dq_status dq_spmv(dq_system* sys, int32_t traj, const double* x, double* y)
{
    if (!sys || !x || !y) return fail(DQ_ERR_INVALID_ARG, "null argument");
    if (!sys->have_h || !sys->have_disorder) return fail(DQ_ERR_STATE, "Hamiltonian and disorder required");
    if (traj < 0 || traj >= sys->B) return fail(DQ_ERR_INVALID_ARG, "trajectory out of range");
    const Dev& d = sys->d;
    const int n = sys->n, Bp = d.Bp;
    const size_t nB = (size_t)n * Bp;
    const int na = sys->nact_slot[sys->slot[traj]];
    traj = sys->slot[traj];
    // rows beyond n_active must stay 0 in every basis slot
    if (d.mixed) {
        std::vector<float2> h(nB, make_float2(0.f, 0.f));
        for (int i = 0; i < na; ++i) h[(size_t)i * Bp + traj] = make_float2((float)x[2 * i], (float)x[2 * i + 1]);
        HIP_TRY(hipMemcpyAsync(d.V, h.data(), sizeof(float2) * nB, hipMemcpyHostToDevice, sys->stream));
        HIP_TRY(dq::launch_spmm(d, 0, 0, sys->stream));
        HIP_TRY(hipStreamSynchronize(sys->stream));
        HIP_TRY(hipMemcpy(h.data(), (float2*)d.V + d.vs, sizeof(float2) * nB, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) { y[2 * i] = h[(size_t)i * Bp + traj].x; y[2 * i + 1] = h[(size_t)i * Bp + traj].y; }
    } else {
        std::vector<double2> h(nB, make_double2(0.0, 0.0));
        for (int i = 0; i < na; ++i) h[(size_t)i * Bp + traj] = make_double2(x[2 * i], x[2 * i + 1]);
        HIP_TRY(hipMemcpyAsync(d.V, h.data(), sizeof(double2) * nB, hipMemcpyHostToDevice, sys->stream));
        HIP_TRY(dq::launch_spmm(d, 0, 0, sys->stream));
        HIP_TRY(hipStreamSynchronize(sys->stream));
        HIP_TRY(hipMemcpy(h.data(), (double2*)d.V + d.vs, sizeof(double2) * nB, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) { y[2 * i] = h[(size_t)i * Bp + traj].x; y[2 * i + 1] = h[(size_t)i * Bp + traj].y; }
    }
    return DQ_OK;
}

dq_status dq_pade_tbar(int32_t m, int32_t count, const double* alpha, const double* beta, double theta,
                       double* E, int32_t* s)
{
    if (m < 2 || m > dq::kMaxM || count <= 0 || !alpha || !beta || !E || !s)
        return fail(DQ_ERR_INVALID_ARG, "bad arguments");
    const int n = m + 2;
    double *da = nullptr, *db = nullptr;
    double2* dE = nullptr;
    int* ds = nullptr;
    hipError_t e = hipSuccess;
    auto cleanup = [&]() { hipFree(da); hipFree(db); hipFree(dE); hipFree(ds); };
    if ((e = hipMalloc(&da, sizeof(double) * m * count)) != hipSuccess ||
        (e = hipMalloc(&db, sizeof(double) * m * count)) != hipSuccess ||
        (e = hipMalloc(&dE, sizeof(double2) * n * n * count)) != hipSuccess ||
        (e = hipMalloc(&ds, sizeof(int) * count)) != hipSuccess) {
        cleanup();
        return fail(DQ_ERR_OOM, hipGetErrorString(e));
    }
    hipMemcpy(da, alpha, sizeof(double) * m * count, hipMemcpyHostToDevice);
    hipMemcpy(db, beta, sizeof(double) * m * count, hipMemcpyHostToDevice);
    hipMemset(dE, 0, sizeof(double2) * n * n * count);
    e = dq::launch_pade_tbar(m, count, da, db, theta, dE, ds, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(E, dE, sizeof(double2) * n * n * count, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(s, ds, sizeof(int) * count, hipMemcpyDeviceToHost);
    cleanup();
    if (e != hipSuccess) return fail(DQ_ERR_HIP, hipGetErrorString(e));
    for (int c = 0; c < count; ++c)
        if (s[c] < 0) return fail(DQ_ERR_SCALING_OVERFLOW, "Pade scaling exponent s >= 50");
    return DQ_OK;
}

dq_status dq_fft(int32_t n, int32_t batch, double* data, int32_t sign)
{
    if (n < 1 || (n & (n - 1)) || n > 65536 || batch < 1 || !data) return fail(DQ_ERR_INVALID_ARG, "bad FFT arguments");
    double2 *a = nullptr, *b = nullptr;
    const size_t bytes = sizeof(double2) * (size_t)n * batch;
    hipError_t e;
    if ((e = hipMalloc(&a, bytes)) != hipSuccess || (e = hipMalloc(&b, bytes)) != hipSuccess) {
        hipFree(a); hipFree(b);
        return fail(DQ_ERR_OOM, hipGetErrorString(e));
    }
    e = hipMemcpy(a, data, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = dq::launch_fft(n, batch, a, b, sign, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(data, a, bytes, hipMemcpyDeviceToHost);
    hipFree(a); hipFree(b);
    if (e != hipSuccess) return fail(DQ_ERR_HIP, hipGetErrorString(e));
    return DQ_OK;
}

}  // extern "C"

// Spectrum of C(n) = (wts ? sum_b wts[b] C_b(n) : batch mean) times scale / T:
// window, zero padding to nfft, inverse-direction FFT on the GPU
// (spectraClass.cu:575-750).  wts_host: per caller trajectory, or null.
// work: null (the buffers are allocated here), or 2 * nfft complex doubles of
// device memory for the windowed C(n) and the FFT's second buffer (no
// weights then): every element is written before it is read.
static dq_status spectrum_impl(dq_system* sys, double sigma_fs, double scale_mu2, const double* wts_host,
                               int32_t nfft, double* omega, double* spec, double2* work = nullptr)
{
    if (!sys || !omega || !spec) return fail(DQ_ERR_INVALID_ARG, "null argument");
    if (!sys->have_bra) return fail(DQ_ERR_STATE, "no correlation function (bra not set)");
    if (nfft < 2 || (nfft & (nfft - 1))) return fail(DQ_ERR_INVALID_ARG, "nfft must be a power of two");
    const int64_t T = std::min(sys->step, sys->max_steps);
    if (T < 1 || T + 1 > nfft) return fail(DQ_ERR_INVALID_ARG, "need 1 <= steps < nfft");
    std::vector<double> wslot;
    if (wts_host) {
        wslot.assign(sys->d.Bp, 0.0);
        for (int b = 0; b < sys->B; ++b) {
            if (!std::isfinite(wts_host[b])) return fail(DQ_ERR_NONFINITE, "non-finite spectrum weight");
            wslot[sys->slot[b]] = wts_host[b];
        }
    }
    double2 *a = nullptr, *b = nullptr;
    double* w = nullptr;
    const size_t bytes = sizeof(double2) * (size_t)nfft;
    hipError_t e = hipSuccess;
    const bool own = work == nullptr;
    if (!own) {
        if (wts_host) return fail(DQ_ERR_INVALID_ARG, "weights need allocated buffers");
        a = work;
        b = work + nfft;
    } else if ((e = hipMalloc(&a, bytes)) != hipSuccess || (e = hipMalloc(&b, bytes)) != hipSuccess ||
               (wts_host && (e = hipMalloc(&w, sizeof(double) * wslot.size())) != hipSuccess)) {
        hipFree(a); hipFree(b); hipFree(w);
        return fail(DQ_ERR_OOM, hipGetErrorString(e));
    }
    if (w) e = hipMemcpyAsync(w, wslot.data(), sizeof(double) * wslot.size(), hipMemcpyHostToDevice, sys->stream);
    if (e == hipSuccess) e = dq::launch_spectrum_prep(sys->d, T, sys->cfg.dt_fs, sigma_fs, nfft, a, w, sys->stream);
    if (e == hipSuccess) e = dq::launch_fft(nfft, 1, a, b, +1, sys->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(sys->stream);
    if (e == hipSuccess) e = hipMemcpy(spec, a, bytes, hipMemcpyDeviceToHost);
    if (own) { hipFree(a); hipFree(b); hipFree(w); }
    if (e != hipSuccess) return fail(DQ_ERR_HIP, hipGetErrorString(e));
    // spectraClass.cu:147-151 (true lightspeed on the frequency axis) and
    // the |mu|^2 / T scaling of spectraClass.cu:641-686
    const double inv_fs_to_inv_cm = 1. / (299792458. * 100. * 1.e-15);
    const double scale = scale_mu2 / (double)T;
    for (int k = 0; k < nfft; ++k) {
        const int kk = k > nfft / 2 ? k - nfft : k;
        omega[k] = inv_fs_to_inv_cm * (double)kk / (sys->cfg.dt_fs * (double)nfft);
        spec[2 * k] *= scale;
        spec[2 * k + 1] *= scale;
    }
    return DQ_OK;
}

extern "C" {

dq_status dq_spectrum(dq_system* sys, double sigma_fs, double mu2, int32_t nfft, double* omega, double* spec)
{
    return spectrum_impl(sys, sigma_fs, mu2, nullptr, nfft, omega, spec);
}

// dq_spectrum in the library's fixed device workspace (dq_noalloc.hip): no
// device memory is allocated or freed.  Calls that overlap on one GPU share
// that workspace and race.
dq_status dq_spectrum_noalloc(dq_system* sys, double sigma_fs, double mu2, int32_t nfft, double* omega, double* spec)
{
    if (!sys || !omega || !spec) return fail(DQ_ERR_INVALID_ARG, "null argument");
    if (nfft > dq::kNoAllocMaxNfft)
        return fail(DQ_ERR_INVALID_ARG, "nfft exceeds the fixed workspace of dq_spectrum_noalloc (131072)");
    double2* work = nullptr;
    HIP_TRY(dq::noalloc_scratch(&work));
    return spectrum_impl(sys, sigma_fs, mu2, nullptr, nfft, omega, spec, work);
}

dq_status dq_spectrum_weighted(dq_system* sys, double sigma_fs, const double* weight, int32_t nfft, double* omega,
                               double* spec)
{
    if (!weight) return fail(DQ_ERR_INVALID_ARG, "null weights");
    return spectrum_impl(sys, sigma_fs, 1.0, weight, nfft, omega, spec);
}

// Several bras: trajectory b records C_b(n) = <bras[bra_of_traj[b]]|psi_b(n)>
// (the reference runs one ensemble per dipole direction, each with its own
// bra, spectraClass.cu:437-461, 699-744)
dq_status dq_system_set_bras(dq_system* sys, int32_t nbra, const double* bras, const int32_t* bra_of_traj)
{
    if (!sys || !bras || !bra_of_traj) return fail(DQ_ERR_INVALID_ARG, "null argument");
    if (nbra < 1) return fail(DQ_ERR_INVALID_ARG, "nbra must be >= 1");
    Dev& d = sys->d;
    const int n = sys->n;
    for (size_t q = 0; q < (size_t)2 * n * nbra; ++q)
        if (!std::isfinite(bras[q])) return fail(DQ_ERR_NONFINITE, "non-finite bra");
    std::vector<int32_t> idx(d.Bp, 0);
    for (int b = 0; b < sys->B; ++b) {
        if (bra_of_traj[b] < 0 || bra_of_traj[b] >= nbra) return fail(DQ_ERR_INVALID_ARG, "bra index out of range");
        idx[sys->slot[b]] = bra_of_traj[b];
    }
    sys->drop_graph();
    HIP_TRY(hipStreamSynchronize(sys->stream));
    Pending pend(sys);
    double2* br = nullptr;
    int32_t* bi = nullptr;
    double2* corr = nullptr;
    DQ_TRY(pend.alloc(&br, (size_t)n * nbra));
    DQ_TRY(pend.alloc(&bi, (size_t)d.Bp));
    if (!d.corr) DQ_TRY(pend.alloc(&corr, (size_t)(sys->max_steps + 1) * d.Bp));
    HIP_TRY(hipMemcpyAsync(br, bras, sizeof(double2) * n * nbra, hipMemcpyHostToDevice, sys->stream));
    HIP_TRY(hipMemcpyAsync(bi, idx.data(), sizeof(int32_t) * d.Bp, hipMemcpyHostToDevice, sys->stream));
    HIP_TRY(hipStreamSynchronize(sys->stream));
    pend.keep();
    sys->release(*(double2**)&d.bra);
    sys->release(*(int32_t**)&d.bra_idx);
    d.bra = br;
    d.bra_idx = bi;
    if (corr) d.corr = corr;
    sys->have_bra = true;
    return DQ_OK;
}


// ---------------------------------------------------------------------------
// adjoint
// ---------------------------------------------------------------------------
dq_status dq_adjoint(dq_system* sys, int64_t nsteps, int32_t ckpt, const double* w_pop, double w_spec,
                     const double* corr_target, const double* corr_weight, double* loss, double* grad_e,
                     double* grad_g)
{
    DQ_TRY(check_ready(sys));
    Dev& d = sys->d;
    if (d.plain || d.mixed) return fail(DQ_ERR_INVALID_ARG, "the adjoint runs on batched FP64 systems (force_batched)");
    if (sys->tk.p > 0) return fail(DQ_ERR_INVALID_ARG, "the adjoint differentiates the Lanczos propagator, not Taylor-Krylov");
    if (nsteps < 1 || ckpt < 1) return fail(DQ_ERR_INVALID_ARG, "nsteps and ckpt must be >= 1");
    if (sys->step != 0) return fail(DQ_ERR_STATE, "the adjoint starts from step 0 (reset_disorder and set_state)");
    if (!loss || !grad_e || !grad_g) return fail(DQ_ERR_INVALID_ARG, "null output");
    if (sys->max_steps < nsteps) return fail(DQ_ERR_INVALID_ARG, "observables max_steps < nsteps");
    const bool spec = sys->have_bra && w_spec != 0.0;
    if (w_spec != 0.0 && !sys->have_bra) return fail(DQ_ERR_STATE, "spectral loss needs a bra");
    if (w_pop && (sys->every <= 0 || sys->site_mode)) return fail(DQ_ERR_STATE, "population loss needs group observables");
    hipStream_t st = sys->stream;
    const int n = sys->n, Bp = d.Bp, m = sys->m, K = sys->K;
    const size_t nB = (size_t)n * Bp;
    ckpt = (int32_t)std::min<int64_t>(ckpt, nsteps);
    const int ncyc = ckpt * K;
    const int64_t nck = (nsteps + ckpt - 1) / ckpt;
    // the forward Pade kernel accepts s < 50 (padeClass.cu:618-623); the
    // reverse tapes every squaring, so its scratch is sized for the same bound
    const int max_s = 49;
    std::vector<void*> tmp;
    auto get = [&](void** p, size_t bytes) -> dq_status {
        hipError_t e = hipMalloc(p, bytes > 0 ? bytes : 16);
        if (e != hipSuccess) return fail(DQ_ERR_OOM, std::string("adjoint hipMalloc: ") + hipGetErrorString(e));
        tmp.push_back(*p);
        hipMemsetAsync(*p, 0, bytes > 0 ? bytes : 16, st);
        return DQ_OK;
    };
    // every exit, error paths included, frees the adjoint's buffers
    struct Scratch {
        hipStream_t st;
        std::vector<void*>& tmp;
        ~Scratch()
        {
            hipStreamSynchronize(st);
            for (void* p : tmp) hipFree(p);
            tmp.clear();
        }
    } scratch_guard{st, tmp};
    dq::AdjDev a{};
    a.m = m; a.K = K; a.nB = nB; a.vs = d.vs; a.Bp = Bp; a.max_s = max_s;
    double2 *ck_psi, *ad_pbar;
    double *ck_de, *ck_nrm;
    dq_status rc = DQ_OK;
#define AGET(p, bytes) do { if ((rc = get((void**)(p), (bytes))) != DQ_OK) return rc; } while (0)
    AGET(&ck_psi, sizeof(double2) * nB * nck);
    AGET(&ck_de, sizeof(double) * nB * nck);
    AGET(&ck_nrm, sizeof(double) * Bp * nck);
    AGET(&a.tapeV, sizeof(double2) * d.vs * (m + 1) * ncyc);
    AGET(&a.sc, sizeof(double) * Bp * (m + 1) * ncyc);
    AGET(&a.alpha, sizeof(double) * Bp * m * ncyc);
    AGET(&a.beta, sizeof(double) * Bp * m * ncyc);
    AGET(&a.nrm2, sizeof(double) * Bp * ncyc);
    AGET(&a.tdiag, sizeof(double) * nB * ckpt);
    AGET(&a.final_psi, sizeof(double2) * nB);
    // the forward run's final noise offsets, diagonal and error sums, put back
    // after the reverse sweep has recomputed earlier segments over them
    double *fin_de, *fin_diag, *fin_err;
    AGET(&fin_de, sizeof(double) * nB);
    AGET(&fin_diag, sizeof(double) * nB);
    AGET(&fin_err, sizeof(double) * Bp);
    AGET(&a.pbar, sizeof(double2) * nB);
    AGET(&a.pbar1, sizeof(double2) * nB);
    AGET(&a.wb, sizeof(double2) * nB);
    for (int q = 0; q < 3; ++q) AGET(&a.qb[q], sizeof(double2) * nB);
    AGET(&a.grad_e, sizeof(double) * nB);
    AGET(&a.gpart, sizeof(double) * Bp * d.ntile_s);
    AGET(&a.y, sizeof(double2) * Bp * m);
    AGET(&a.abar, sizeof(double) * Bp * m);
    AGET(&a.bbar, sizeof(double) * Bp * m);
    AGET(&a.r2, sizeof(double) * Bp);
    AGET(&a.r3, sizeof(double) * Bp * m);
    AGET(&a.cbar, sizeof(double2) * Bp * (nsteps + 1));
    a.pade_stride = (size_t)(16 + max_s + 1) * (dq::kMaxM + 2) * (dq::kMaxM + 2);
    AGET(&a.pade_scratch, sizeof(double2) * a.pade_stride * Bp);
    (void)ad_pbar;
    // a cycle of the current segment writes its basis, scalars and diagonal
    // into the tape
    auto tape_dev = [&](int c) {
        Dev dc = d;
        dc.V = a.tapeV + (size_t)c * (m + 1) * d.vs;
        dc.sc = a.sc + (size_t)c * (m + 1) * Bp;
        dc.alpha = a.alpha + (size_t)c * m * Bp;
        dc.beta = a.beta + (size_t)c * m * Bp;
        dc.diag = a.tdiag + (size_t)(c / K) * nB;
        return dc;
    };
    const bool obs = dq::obs_on(d);
    // ---------------- forward with checkpoints ----------------
    // the last segment is taped on the way, so the reverse does not recompute it
    const int64_t last_s0 = (nck - 1) * ckpt;
    for (int64_t step = 0; step < nsteps; ++step) {
        if (step % ckpt == 0) {
            const int64_t k = step / ckpt;
            HIP_TRY(hipMemcpyAsync(ck_psi + nB * k, d.psi, sizeof(double2) * nB, hipMemcpyDeviceToDevice, st));
            HIP_TRY(hipMemcpyAsync(ck_de + nB * k, d.de, sizeof(double) * nB, hipMemcpyDeviceToDevice, st));
            HIP_TRY(hipMemcpyAsync(ck_nrm + (size_t)Bp * k, d.nrm2, sizeof(double) * Bp, hipMemcpyDeviceToDevice, st));
        }
        if (step < last_s0) {
            HIP_TRY(enqueue_step(sys));
            continue;
        }
        for (int k = 0; k < K; ++k) {
            const int c = (int)(step - last_s0) * K + k;
            HIP_TRY(enqueue_cycle_dev(sys, tape_dev(c), k == 0, k == K - 1, obs));
            HIP_TRY(hipMemcpyAsync(a.nrm2 + (size_t)c * Bp, d.nrm2, sizeof(double) * Bp, hipMemcpyDeviceToDevice, st));
        }
    }
    HIP_TRY(enqueue_finalize(sys));
    HIP_TRY(hipMemcpyAsync(a.final_psi, d.psi, sizeof(double2) * nB, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(fin_de, d.de, sizeof(double) * nB, hipMemcpyDeviceToDevice, st));
    // the taped last segment wrote its diagonals into the tape: the final
    // step's is the last of them
    HIP_TRY(hipMemcpyAsync(fin_diag, a.tdiag + (size_t)(nsteps - 1 - last_s0) * nB, sizeof(double) * nB,
                           hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(fin_err, d.err_sum, sizeof(double) * Bp, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
    sys->step = nsteps;
    DQ_TRY(dq_sync(sys));
    // ---------------- loss and its adjoint at T (host) ----------------
    std::vector<double2> hpsi(nB), hbar(nB, make_double2(0.0, 0.0));
    HIP_TRY(hipMemcpy(hpsi.data(), d.psi, sizeof(double2) * nB, hipMemcpyDeviceToHost));
    std::vector<int32_t> grp(n, -1);
    if (w_pop) HIP_TRY(hipMemcpy(grp.data(), d.group, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    std::vector<double2> bra(n, make_double2(0.0, 0.0));
    std::vector<double2> corr, cb((size_t)(nsteps + 1) * Bp, make_double2(0.0, 0.0));
    if (spec) {
        HIP_TRY(hipMemcpy(bra.data(), d.bra, sizeof(double2) * n, hipMemcpyDeviceToHost));
        corr.resize((size_t)(nsteps + 1) * Bp);
        HIP_TRY(hipMemcpy(corr.data(), d.corr, sizeof(double2) * corr.size(), hipMemcpyDeviceToHost));
    }
    for (int b = 0; b < sys->B; ++b) {
        const int sb = sys->slot[b];
        double L = 0.0;
        if (w_pop)
            for (int i = 0; i < n; ++i) {
                const int g = grp[i];
                if (g < 0) continue;
                const double2 p = hpsi[(size_t)i * Bp + sb];
                L += w_pop[g] * (p.x * p.x + p.y * p.y);
                hbar[(size_t)i * Bp + sb] = make_double2(2.0 * w_pop[g] * p.x, 2.0 * w_pop[g] * p.y);
            }
        if (spec)
            for (int64_t k = 0; k <= nsteps; ++k) {
                const double2 c = corr[(size_t)k * Bp + sb];
                const double tr = corr_target ? corr_target[2 * k] : 0.0;
                const double ti = corr_target ? corr_target[2 * k + 1] : 0.0;
                const double dr = c.x - tr, di = c.y - ti;
                const double ws = w_spec * (corr_weight ? corr_weight[k] : 1.0);
                L += ws * (dr * dr + di * di);
                cb[(size_t)k * Bp + sb] = make_double2(2.0 * ws * dr, 2.0 * ws * di);
            }
        loss[b] = L;
    }
    HIP_TRY(hipMemcpyAsync(a.pbar, hbar.data(), sizeof(double2) * nB, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(a.cbar, cb.data(), sizeof(double2) * cb.size(), hipMemcpyHostToDevice, st));
    if (spec) HIP_TRY(dq::adj_add_corr(d, a, nsteps, a.pbar, st));
    // ---------------- reverse over segments ----------------
    for (int64_t seg = nck - 1; seg >= 0; --seg) {
        const int64_t s0 = seg * ckpt, s1 = std::min<int64_t>(s0 + ckpt, nsteps);
        const int nc = (int)(s1 - s0) * K;
        if (seg < nck - 1) {             // the last segment was taped by the forward pass
            HIP_TRY(hipMemcpyAsync(d.psi, ck_psi + nB * seg, sizeof(double2) * nB, hipMemcpyDeviceToDevice, st));
            HIP_TRY(hipMemcpyAsync(d.de, ck_de + nB * seg, sizeof(double) * nB, hipMemcpyDeviceToDevice, st));
            HIP_TRY(hipMemcpyAsync(d.nrm2, ck_nrm + (size_t)Bp * seg, sizeof(double) * Bp, hipMemcpyDeviceToDevice, st));
            int64_t s0v = s0;
            HIP_TRY(hipMemcpyAsync(d.step, &s0v, sizeof(int64_t), hipMemcpyHostToDevice, st));
            HIP_TRY(hipStreamSynchronize(st));
            for (int c = 0; c < nc; ++c) {
                const int k = c % K;
                HIP_TRY(enqueue_cycle_dev(sys, tape_dev(c), k == 0, k == K - 1, false));
                HIP_TRY(hipMemcpyAsync(a.nrm2 + (size_t)c * Bp, d.nrm2, sizeof(double) * Bp, hipMemcpyDeviceToDevice, st));
            }
            {
                Dev dc = d;
                dc.V = a.final_psi;
                HIP_TRY(dq::launch_prologue(dc, 1, 0, 0, 1, st));
            }
        }
        a.last_c = nc - 1;
        for (int c = nc - 1; c >= 0; --c) {
            HIP_TRY(dq::adj_renorm(d, a, c, st));
            HIP_TRY(dq::adj_pade(d, a, c, st));
            HIP_TRY(dq::adj_lanczos(d, a, c, st));
            HIP_TRY(hipMemcpyAsync(a.pbar, a.qbar_slot(0), sizeof(double2) * nB, hipMemcpyDeviceToDevice, st));
            if (c % K == 0 && spec) HIP_TRY(dq::adj_add_corr(d, a, s0 + c / K, a.pbar, st));
        }
        HIP_TRY(hipStreamSynchronize(st));
    }
    DQ_TRY(dq_sync(sys));
    // ---------------- gradients ----------------
    std::vector<double> ge(nB), gp((size_t)Bp * d.ntile_s);
    HIP_TRY(hipMemcpy(ge.data(), a.grad_e, sizeof(double) * nB, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(gp.data(), a.gpart, sizeof(double) * gp.size(), hipMemcpyDeviceToHost));
    for (int b = 0; b < sys->B; ++b) {
        const int sb = sys->slot[b];
        for (int i = 0; i < n; ++i) grad_e[(size_t)b * n + i] = ge[(size_t)i * Bp + sb];
        double g = 0.0;
        for (int t = 0; t < d.ntile_s; ++t) g += gp[(size_t)t * Bp + sb];
        grad_g[b] = g;
    }
    // leave the system at the final state of the forward run: psi, noise
    // offsets, diagonal, error sums and step counter
    HIP_TRY(hipMemcpyAsync(d.psi, hpsi.data(), sizeof(double2) * nB, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(d.de, fin_de, sizeof(double) * nB, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(d.diag, fin_diag, sizeof(double) * nB, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(d.err_sum, fin_err, sizeof(double) * Bp, hipMemcpyDeviceToDevice, st));
    int64_t sT = nsteps;
    HIP_TRY(hipMemcpyAsync(d.step, &sT, sizeof(int64_t), hipMemcpyHostToDevice, st));
    HIP_TRY(dq::launch_set_unit_norm(d, st));
#undef AGET
    return DQ_OK;
}

// ---------------------------------------------------------------------------
// Launch tuning and the autotuner
// ---------------------------------------------------------------------------
static bool tuning_valid(const dq_system* sys, const dq_tuning& t, std::string* why)
{
    const Dev& d = sys->d;
    if (t.spmm_rows_per_wave != 2 && t.spmm_rows_per_wave != 4 && t.spmm_rows_per_wave != 8 &&
        t.spmm_rows_per_wave != 16) {
        *why = "spmm_rows_per_wave must be 2, 4, 8 or 16";
        return false;
    }
    if (t.spmm_rows_per_wave == 16 && !d.mixed) {
        *why = "spmm_rows_per_wave = 16 (the matrix-core SpMM) is for mixed precision";
        return false;
    }
    if (t.spmm_rows_per_wave == 16) {
        // 32-bit offsets (bounded above by nnz + 15 padding slots per 16-row block)
        const size_t nblk16 = (size_t)(sys->n + 15) / 16;
        const size_t nnz_max = std::max(sys->col.size(), sys->tk.p > 0 ? sys->tk.col_h.size() : (size_t)0);
        if (!mfma_offsets_fit(sys->n, d.Bp, nnz_max + 15 * nblk16)) {
            *why = "spmm_rows_per_wave = 16 needs n * padded batch * 8 and 64 * (nnz + 15 n / 16) below 2^31 bytes";
            return false;
        }
        // matrix-core SpMM: 1, 2 or 4 trajectories per lane = 16, 32, 64 per wave
        if (t.spmm_traj_per_lane != 1 && t.spmm_traj_per_lane != 2 && t.spmm_traj_per_lane != 4) {
            *why = "with spmm_rows_per_wave = 16, spmm_traj_per_lane must be 1, 2 or 4";
            return false;
        }
        if (t.spmm_phases != 1 || t.spmm_pipeline != 0) {
            *why = "with spmm_rows_per_wave = 16, spmm_phases must be 1 and spmm_pipeline 0";
            return false;
        }
    } else if (t.spmm_traj_per_lane != 1 && t.spmm_traj_per_lane != 2) {
        *why = "spmm_traj_per_lane must be 1 or 2";
        return false;
    }
    if (t.spmm_rows_per_wave != 16 && t.spmm_traj_per_lane == 2 &&
        !(d.mixed && d.Bp % 128 == 0)) {
        *why = "spmm_traj_per_lane = 2 needs mixed precision and a padded batch that is a multiple of 128";
        return false;
    }
    const int nslice = t.spmm_rows_per_wave == 16 ? 1 : d.Bp / (64 * t.spmm_traj_per_lane);
    if ((t.spmm_phases != 1 && t.spmm_phases != 2 && t.spmm_phases != 4) || nslice % t.spmm_phases) {
        *why = "spmm_phases must be 1, 2 or 4 and divide the number of trajectory slices";
        return false;
    }
    if (t.tile_rows != 32 && t.tile_rows != 64 && t.tile_rows != 128 && t.tile_rows != 256) {
        *why = "tile_rows must be 32, 64, 128 or 256";
        return false;
    }
    if (t.spmm_pipeline != 0 && t.spmm_pipeline != 1) {
        *why = "spmm_pipeline must be 0 or 1";
        return false;
    }
    if (t.spmm_exact_sums != 0 && t.spmm_exact_sums != 1) {
        *why = "spmm_exact_sums must be 0 or 1";
        return false;
    }
    const int ns = t.concurrent_slices < 1 ? 1 : t.concurrent_slices;
    if (ns != 1 && ns != 2 && ns != 4) {
        *why = "concurrent_slices must be 1, 2 or 4";
        return false;
    }
    if (d.Bp % (64 * ns) != 0) {
        *why = "concurrent_slices needs slices of a multiple of 64 trajectory slots";
        return false;
    }
    if (t.spmm_window_rows < 0) {
        *why = "spmm_window_rows must be >= 0";
        return false;
    }
    if (t.spmm_window_rows > 0) {
        if (!d.mixed) {
            *why = "the ring SpMM (spmm_window_rows > 0) is for mixed precision";
            return false;
        }
        if (t.spmm_window_rows % dq::kRingChunk != 0 || t.spmm_window_rows > dq::kRingMaxRows) {
            *why = "spmm_window_rows must be a multiple of 128, at most 4096";
            return false;
        }
        if ((uint64_t)sys->n * (uint64_t)d.Bp * 8 >= ((uint64_t)1 << 31)) {
            *why = "the ring SpMM addresses a Krylov vector with 32-bit offsets: n * padded batch * 8 "
                   "must stay below 2^31 bytes";
            return false;
        }
        if (t.spmm_window_waves != 0 && t.spmm_window_waves != 4 && t.spmm_window_waves != 8 &&
            t.spmm_window_waves != 16) {
            *why = "spmm_window_waves must be 4, 8 or 16 (0: 16)";
            return false;
        }
    }
    return true;
}

static void get_tuning(const dq_system* sys, dq_tuning* t)
{
    std::memset(t, 0, sizeof(*t));
    t->spmm_rows_per_wave = sys->d.rb;
    t->spmm_traj_per_lane = sys->d.spmm_tpl;
    t->spmm_phases = sys->d.spmm_phases;
    t->tile_rows = sys->d.plain ? 0 : sys->d.rt_e;
    t->spmm_pipeline = sys->d.spmm_pipe;
    t->spmm_exact_sums = sys->d.spmm_exact;
    t->concurrent_slices = sys->slices;
    t->spmm_window_rows = sys->ring_rows;
    t->spmm_window_waves = sys->ring_rows > 0 ? sys->ring_waves : 0;
}

// Applies a validated tuning; the row blocks and ring lists it needs are
// built before anything is replaced, so a failure (a band too wide for the
// ring) leaves the previous choice in place.
static dq_status apply_tuning(dq_system* sys, const dq_tuning& t)
{
    Dev& d = sys->d;
    const int T = t.spmm_window_rows;
    RingLists nr;
    const bool new_ring = T > 0 && sys->have_h && !sys->ring.item;
    if (new_ring) DQ_TRY(make_ring_lists(sys, &nr));
    if (t.spmm_rows_per_wave != d.rb) {
        if (sys->have_h) {
            dq_status st = build_row_blocks(sys, t.spmm_rows_per_wave);
            if (st != DQ_OK) {
                release_ring(sys, &nr);
                return st;
            }
        } else {
            d.rb = t.spmm_rows_per_wave;
        }
    }
    // the lists follow the Hamiltonian only while the ring SpMM is chosen
    // (set_hamiltonian rebuilds them then), so they go when it is not
    if (T == 0) release_ring(sys, &sys->ring);
    else if (new_ring) sys->ring = nr;
    sys->ring_rows = T;
    sys->ring_waves = t.spmm_window_waves > 0 ? t.spmm_window_waves : 16;
    d.spmm_tpl = t.spmm_traj_per_lane;
    d.spmm_phases = t.spmm_phases;
    d.spmm_pipe = t.spmm_pipeline;
    d.spmm_exact = t.spmm_exact_sums;
    sys->slices = t.concurrent_slices < 1 ? 1 : t.concurrent_slices;
    d.rt_e = t.tile_rows;
    d.ntile_e = (d.n + d.rt_e - 1) / d.rt_e;
    attach_ring(sys);
    return DQ_OK;
}

// one Lanczos cycle without the Pade error sum, the renormalisation and the
// observables: once the noise is set up, the noise prologue, then SpMMs,
// reductions, scalars, updates and the combine.  The caller saves and restores psi, the noise
// offsets and the diagonal around it.
static bool tuning_with_noise(const dq_system* sys)
{
    return sys->noise_ready() && sys->have_disorder;
}

static hipError_t enqueue_tuning_cycle(dq_system* sys, const Dev& d, hipStream_t st)
{
    hipError_t e;
    if (tuning_with_noise(sys) && (e = dq::launch_prologue(d, 0, 1, 0, 1, st)) != hipSuccess) return e;
    for (int j = 0; j < d.m; ++j) {
        if ((e = dq::launch_spmm(d, j, 0, st)) != hipSuccess) return e;
        if (!d.plain && !dq::spmm_fuses_gram(d) && (e = dq::launch_reduce_all(d, j, st)) != hipSuccess) return e;
        if ((e = dq::launch_scalars_h(d, j, st)) != hipSuccess) return e;
        if ((e = dq::launch_update(d, j, st)) != hipSuccess) return e;
        if (d.plain && (e = dq::launch_reduce(d, 2 * (j + 1) + 1, d.ntile_e, st)) != hipSuccess) return e;
    }
    if ((e = dq::launch_spmm(d, d.m, 1, st)) != hipSuccess) return e;
    return dq::launch_combine(d, st);
}

static hipError_t enqueue_tuning_cycle(dq_system* sys)
{
    return enqueue_tuning_cycle(sys, sys->d, sys->stream);
}

// the tuning cycle over `ns` concurrent slices, captured as one graph
static hipError_t capture_tuning_slices(dq_system* sys, int ns, hipGraphExec_t* out)
{
    const Dev& d = sys->d;
    hipGraph_t g;
    hipError_t e = hipStreamBeginCapture(sys->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    e = hipEventRecord(sys->ev_fork, sys->stream);
    for (int h = 1; h < ns && e == hipSuccess; ++h) e = hipStreamWaitEvent(sys->side[h - 1], sys->ev_fork, 0);
    for (int h = 0; h < ns && e == hipSuccess; ++h) {
        Dev dh = d;
        dh.sn = d.Bp / ns;
        dh.s0 = h * dh.sn;
        e = enqueue_tuning_cycle(sys, dh, h ? sys->side[h - 1] : sys->stream);
    }
    for (int h = 1; h < ns && e == hipSuccess; ++h) {
        e = hipEventRecord(sys->ev_join[h - 1], sys->side[h - 1]);
        if (e == hipSuccess) e = hipStreamWaitEvent(sys->stream, sys->ev_join[h - 1], 0);
    }
    hipError_t e2 = hipStreamEndCapture(sys->stream, &g);
    if (e != hipSuccess) return e;
    if (e2 != hipSuccess) return e2;
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    return e;
}

static std::string tuning_key(const dq_system* sys)
{
    hipDeviceProp_t prop;
    int dev = 0;
    std::string arch = "unknown";
    int cus = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
        arch = prop.gcnArchName;
        cus = prop.multiProcessorCount;
        for (char& c : arch) if (c == ' ') c = '_';
    }
    char buf[256];
    std::snprintf(buf, sizeof(buf), "%s cu=%d n=%d nnz=%lld bp=%d m=%d prec=%s exact=%d", arch.c_str(), cus, sys->n,
                  (long long)sys->col.size(), sys->Bp, sys->m, sys->d.mixed ? "mixed" : "fp64", sys->d.spmm_exact);
    return buf;
}

dq_status dq_get_tuning(dq_system* sys, dq_tuning* out)
{
    if (!sys || !out) return fail(DQ_ERR_INVALID_ARG, "null argument");
    get_tuning(sys, out);
    return DQ_OK;
}

dq_status dq_set_tuning(dq_system* sys, const dq_tuning* t)
{
    if (!sys || !t) return fail(DQ_ERR_INVALID_ARG, "null argument");
    if (sys->d.plain) return DQ_OK;
    std::string why;
    if (!tuning_valid(sys, *t, &why)) return fail(DQ_ERR_INVALID_ARG, why);
    sys->drop_graph();
    HIP_TRY(hipStreamSynchronize(sys->stream));
    return apply_tuning(sys, *t);
}

dq_status dq_autotune(dq_system* sys, int32_t reps, const char* cache_path, dq_tuning* chosen)
{
    if (!sys) return fail(DQ_ERR_INVALID_ARG, "null system");
    if (!sys->have_h) return fail(DQ_ERR_STATE, "autotune needs set_hamiltonian first");
    if (reps < 1) return fail(DQ_ERR_INVALID_ARG, "reps must be >= 1");
    dq_tuning best;
    get_tuning(sys, &best);
    if (sys->d.plain) {
        if (chosen) *chosen = best;
        return DQ_OK;
    }
    const std::string key = tuning_key(sys);
    if (cache_path) {
        if (FILE* f = std::fopen(cache_path, "r")) {
            char line[512];
            bool hit = false;
            while (!hit && std::fgets(line, sizeof(line), f)) {
                if (std::strncmp(line, key.c_str(), key.size()) != 0 || line[key.size()] != ' ') continue;
                dq_tuning t;
                std::memset(&t, 0, sizeof(t));
                std::string why;
                if (std::sscanf(line + key.size(), "%d %d %d %d %d %d %d %d %d", &t.spmm_rows_per_wave,
                                &t.spmm_traj_per_lane, &t.spmm_phases, &t.tile_rows, &t.spmm_pipeline,
                                &t.concurrent_slices, &t.spmm_window_rows, &t.spmm_window_waves,
                                &t.spmm_exact_sums) == 9 &&
                    tuning_valid(sys, t, &why)) {
                    best = t;
                    hit = true;
                }
            }
            std::fclose(f);
            if (hit) {
                DQ_TRY(dq_set_tuning(sys, &best));
                if (chosen) *chosen = best;
                return DQ_OK;
            }
        }
    }
    Dev& d = sys->d;
    sys->drop_graph();
    // the trial launches stay out of the kernel timings, and the captured
    // trial graphs hold no profiling events
    struct ProfilingOff {
        dq_system* s;
        bool was;
        ~ProfilingOff() { s->profiling = was; }
    } profiling_off{sys, sys->profiling};
    sys->profiling = false;
    HIP_TRY(hipStreamSynchronize(sys->stream));
    const size_t nB = (size_t)d.n * d.Bp;
    // saved state: psi, then the noise offsets and the diagonal
    const size_t diag_bytes = (d.mixed ? sizeof(float) : sizeof(double)) * nB;
    char* saved_raw = nullptr;
    HIP_TRY(hipMalloc(&saved_raw, sizeof(double2) * nB + sizeof(double) * nB + diag_bytes));
    double2* saved = reinterpret_cast<double2*>(saved_raw);
    double* saved_de = reinterpret_cast<double*>(saved_raw + sizeof(double2) * nB);
    void* saved_diag = saved_raw + sizeof(double2) * nB + sizeof(double) * nB;
    dq_tuning orig = best;
    dq_status st = DQ_OK;
    double best_ms = 1e300;
    std::vector<std::pair<float, dq_tuning>> timed_choices;
    hipError_t e = hipMemcpyAsync(saved, d.psi, sizeof(double2) * nB, hipMemcpyDeviceToDevice, sys->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(saved_de, d.de, sizeof(double) * nB, hipMemcpyDeviceToDevice, sys->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(saved_diag, d.diag, diag_bytes, hipMemcpyDeviceToDevice, sys->stream);
    // candidates: the row-blocked SpMMs, then (mixed precision) the ring
    // SpMM, grouped so that each list is built once; the caller's
    // spmm_exact_sums is an accuracy choice and is kept
    const int32_t exact = orig.spmm_exact_sums;
    std::vector<dq_tuning> cands;
    static const int rows[4] = {128, 64, 256, 32};
    for (int R : {4, 2, 8, 16})
        for (int tpl : {2, 1, 4})
            for (int ph : {1, 2, 4})
                for (int cand = 0; cand < 8; ++cand) {
                    // the matrix-core SpMM has no phase or pipeline variants
                    if (R == 16 && (ph != 1 || cand >= 4)) continue;
                    if (R != 16 && tpl == 4) continue;
                    dq_tuning t;
                    std::memset(&t, 0, sizeof(t));
                    t.spmm_rows_per_wave = R;
                    t.spmm_traj_per_lane = tpl;
                    t.spmm_phases = ph;
                    t.tile_rows = rows[cand % 4];
                    t.spmm_pipeline = cand / 4;
                    t.spmm_exact_sums = exact;
                    cands.push_back(t);
                }
    if (d.mixed)
        for (int T : {768, 384, 1536})
            for (int W : {16, 8})
                for (int cand = 0; cand < 4; ++cand) {
                    dq_tuning t;
                    std::memset(&t, 0, sizeof(t));
                    t.spmm_rows_per_wave = 4;
                    t.spmm_traj_per_lane = d.Bp % 128 == 0 ? 2 : 1;
                    t.spmm_phases = 1;
                    t.tile_rows = rows[cand];
                    t.spmm_window_rows = T;
                    t.spmm_window_waves = W;
                    t.spmm_exact_sums = exact;
                    cands.push_back(t);
                }
    for (const dq_tuning& t : cands) {
        if (e != hipSuccess || st != DQ_OK) break;
        std::string why;
        if (!tuning_valid(sys, t, &why)) continue;
        const dq_status sa = apply_tuning(sys, t);
        if (sa == DQ_ERR_INVALID_ARG && t.spmm_window_rows > 0) continue;   // band too wide for the ring
        if ((st = sa) != DQ_OK) break;
        if ((e = enqueue_tuning_cycle(sys)) != hipSuccess) break;       // warm-up
        float ms_min = 1e30f;
        for (int r = 0; r < reps && e == hipSuccess; ++r) {
            float ms = 0.f;
            if ((e = hipEventRecord(sys->ev0, sys->stream)) != hipSuccess) break;
            if ((e = enqueue_tuning_cycle(sys)) != hipSuccess) break;
            if ((e = hipEventRecord(sys->ev1, sys->stream)) != hipSuccess) break;
            if ((e = hipEventSynchronize(sys->ev1)) != hipSuccess) break;
            if ((e = hipEventElapsedTime(&ms, sys->ev0, sys->ev1)) != hipSuccess) break;
            ms_min = std::min(ms_min, ms);
        }
        if (e == hipSuccess) timed_choices.push_back({ms_min, t});
        if (e == hipSuccess && ms_min < best_ms) { best_ms = ms_min; best = t; }
    }
    // concurrent slices: the four fastest kernel choices of one stream, each
    // captured as graphs of 1, 2 and 4 slices (a choice that is fastest alone
    // is not always the one that overlaps best with the other slices)
    if (e == hipSuccess && st == DQ_OK) {
        std::stable_sort(timed_choices.begin(), timed_choices.end(),
                         [](const std::pair<float, dq_tuning>& a, const std::pair<float, dq_tuning>& b) {
                             return a.first < b.first;
                         });
        if (timed_choices.size() > 4) timed_choices.resize(4);
        std::vector<std::pair<float, dq_tuning>> graph_choices;
        for (const auto& cand : timed_choices) {
            if (e != hipSuccess || st != DQ_OK) break;
            if ((st = apply_tuning(sys, cand.second)) != DQ_OK) break;
            for (int ns : {1, 2, 4}) {
                dq_tuning t = cand.second;
                t.concurrent_slices = ns;
                std::string why;
                if (!tuning_valid(sys, t, &why)) continue;
                hipGraphExec_t ge = nullptr;
                if ((e = capture_tuning_slices(sys, ns, &ge)) != hipSuccess) break;
                float ms_min = 1e30f;
                e = hipGraphLaunch(ge, sys->stream);                            // warm-up
                // the candidates differ by a few percent: at least 5 timed replays
                for (int r = 0; r < std::max(reps, 5) && e == hipSuccess; ++r) {
                    float ms = 0.f;
                    if ((e = hipEventRecord(sys->ev0, sys->stream)) != hipSuccess) break;
                    if ((e = hipGraphLaunch(ge, sys->stream)) != hipSuccess) break;
                    if ((e = hipEventRecord(sys->ev1, sys->stream)) != hipSuccess) break;
                    if ((e = hipEventSynchronize(sys->ev1)) != hipSuccess) break;
                    if ((e = hipEventElapsedTime(&ms, sys->ev0, sys->ev1)) != hipSuccess) break;
                    ms_min = std::min(ms_min, ms);
                }
                hipGraphExecDestroy(ge);
                if (e != hipSuccess) break;
                graph_choices.push_back({ms_min, t});
            }
        }
        // the three fastest graphs timed again: the first pass separates them
        // by less than its own spread.  Two rounds over the three (drift
        // between the candidates cancels), each five batches of ten replays
        // back to back; a candidate's figure is its fastest batch.
        std::stable_sort(graph_choices.begin(), graph_choices.end(),
                         [](const std::pair<float, dq_tuning>& a, const std::pair<float, dq_tuning>& b) {
                             return a.first < b.first;
                         });
        if (graph_choices.size() > 3) graph_choices.resize(3);
        std::vector<double> fastest(graph_choices.size(), 1e300);
        for (int round = 0; round < 2; ++round)
            for (size_t c = 0; c < graph_choices.size(); ++c) {
                if (e != hipSuccess || st != DQ_OK) break;
                dq_tuning t = graph_choices[c].second;
                const int ns = t.concurrent_slices;
                t.concurrent_slices = 0;
                if ((st = apply_tuning(sys, t)) != DQ_OK) break;
                hipGraphExec_t ge = nullptr;
                if ((e = capture_tuning_slices(sys, ns, &ge)) != hipSuccess) break;
                e = hipGraphLaunch(ge, sys->stream);                            // warm-up
                for (int r = 0; r < 5 && e == hipSuccess; ++r) {
                    float ms = 0.f;
                    if ((e = hipEventRecord(sys->ev0, sys->stream)) != hipSuccess) break;
                    for (int q = 0; q < 10 && e == hipSuccess; ++q) e = hipGraphLaunch(ge, sys->stream);
                    if (e != hipSuccess) break;
                    if ((e = hipEventRecord(sys->ev1, sys->stream)) != hipSuccess) break;
                    if ((e = hipEventSynchronize(sys->ev1)) != hipSuccess) break;
                    if ((e = hipEventElapsedTime(&ms, sys->ev0, sys->ev1)) != hipSuccess) break;
                    fastest[c] = std::min(fastest[c], (double)ms / 10.0);
                }
                hipGraphExecDestroy(ge);
            }
        double best_graph_ms = 1e300;
        for (size_t c = 0; c < graph_choices.size(); ++c)
            if (fastest[c] < best_graph_ms) { best_graph_ms = fastest[c]; best = graph_choices[c].second; }
    }
    if (e == hipSuccess) e = hipMemcpyAsync(d.psi, saved, sizeof(double2) * nB, hipMemcpyDeviceToDevice, sys->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d.de, saved_de, sizeof(double) * nB, hipMemcpyDeviceToDevice, sys->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d.diag, saved_diag, diag_bytes, hipMemcpyDeviceToDevice, sys->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(sys->stream);
    hipFree(saved_raw);
    if (e != hipSuccess) return fail(DQ_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
    if (st != DQ_OK) { apply_tuning(sys, orig); return st; }
    DQ_TRY(apply_tuning(sys, best));
    if (cache_path) {
        if (FILE* f = std::fopen(cache_path, "a")) {
            std::fprintf(f, "%s %d %d %d %d %d %d %d %d %d\n", key.c_str(), best.spmm_rows_per_wave,
                         best.spmm_traj_per_lane, best.spmm_phases, best.tile_rows, best.spmm_pipeline,
                         best.concurrent_slices, best.spmm_window_rows, best.spmm_window_waves,
                         best.spmm_exact_sums);
            std::fclose(f);
        }
    }
    if (chosen) *chosen = best;
    return DQ_OK;
}

}  // extern "C"
