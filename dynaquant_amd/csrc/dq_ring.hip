// dq_ring.hip -- batched mixed-precision SpMM with the Krylov vector rows
// staged in an LDS ring (lanczosClass.cu:618-661 per trajectory, cusparse
// csrmv on the Lanczos vector).
//
// The Hamiltonian of an aggregate is banded in its site order (a dipole tube
// in ring-major order couples a site only to rings r-3..r+3), so the x rows
// that the rows [r, r + C) reference lie in a short window [lo, hi) that
// slides forward with r.  A workgroup owns 16 trajectories (32 real columns)
// and a range of rows: it holds the window in a ring of 1024 rows per column
// in LDS, reads every x row of its range (plus the band's halo) from HBM
// once, and walks the range in chunks of 128 rows, loading the rows the next
// chunk adds while it multiplies the current one.
//
// The products run on the matrix cores in blocks of 4 rows: the union of the
// 4 rows' columns, grouped in even-aligned pairs (2p, 2p + 1), is the K
// dimension of v_mfma_f32_4x4x1_16b_f32.  Its 16 blocks are 8 groups of 4
// real columns times the two halves of a pair: lanes 0-31 multiply row 2p,
// lanes 32-63 row 2p + 1, and the two halves' sums are added at the end of
// the block.  CBSZ = 3 broadcasts the 4 couplings of one pair and half from
// lanes 4a .. 4a + 3 (resp. 32 + 4a ..), so one coalesced load of 64 floats
// brings the couplings of 8 pairs (ABID = a).  The FP32 couplings are the
// rounded FP64 ones; their bfloat16 low parts go through one
// v_mfma_f32_4x4x4_16b_bf16 per 4 pairs, on the x values already in
// registers, into a sum of their own.  Each group of 4 pairs has its own FP32
// sum (independent chains on the matrix cores); they move into FP64 sums
// after every item of up to 16 pairs (each group's on its own: EXACT), the
// low parts' once per block,
// w = d x + g (J x) is formed in FP64 and rounded once, and the FP64
// partial <x|w> per trajectory goes to part_s (one tile per row range).
//
// A block's pairs come in items of up to 16 (its count padded to a multiple
// of 4): per lane 2 couplings, 4 low parts and the 16 LDS offsets, whose low
// 3 bits (the offsets are multiples of 8) hold the item's pair count and a
// last-item-of-the-block flag.  A wave takes consecutive blocks, so its items
// are consecutive too: the next item's loads are issued while the current
// one multiplies (about 20 matrix-core instructions per item).
#include "dq_device.h"
#include "dq_internal.h"

namespace dq {

constexpr int kRingStride = kRingCap + 1;   // floats per column (odd: the 32 columns a half-wave
                                            // reads in one row fall in 32 distinct banks)
constexpr int kRingTS = 16;                 // trajectories per workgroup

size_t spmm_ring_lds_bytes(int waves, int rows)
{
    // the ring, padded to 8 bytes, the per-wave partial sums, the range's
    // chunk windows and block item indices
    return (((size_t)32 * kRingStride * sizeof(float) + 7) & ~(size_t)7) + (size_t)waves * 2 * 16 * sizeof(double) +
           (size_t)(rows / kRingChunk + 1) * sizeof(int2) + (size_t)(rows / 4 + 2) * sizeof(int);
}

typedef float rf4 __attribute__((ext_vector_type(4)));
typedef short rs4 __attribute__((ext_vector_type(4)));

// two bfloat16 (the high halves of a and b, truncated) in one dword: a low
__device__ __forceinline__ uint32_t ring_bf16_pair(float a, float b)
{
    return __builtin_amdgcn_perm(__float_as_uint(b), __float_as_uint(a), 0x07060302u);
}

// the value of lane ^ 1 (quad_perm [1, 0, 3, 2])
__device__ __forceinline__ float swap_pair(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
}

__device__ __forceinline__ float lds_f32(const float* base, uint32_t byte_off)
{
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + byte_off);
}

// a raw buffer of `bytes` bytes: loads beyond it return 0, stores are dropped
__device__ __forceinline__ __amdgpu_buffer_rsrc_t ring_rsrc(const void* p, int bytes = 0x7fffffff)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, bytes, 0x00020000);
}

// one item: this lane's couplings of pairs 0-7 (a.x) and 8-15 (a.y), its low
// parts (group (lane >> 2) & 3), and the 16 pairs' LDS offsets
struct RingItem {
    float ax, ay;
    uint2 lo;
    uint32_t o[16];
};

// the item's loads: buffer loads at a wave-uniform offset, per-lane parts
// precomputed (the offsets are the same in every lane)
struct RingSrc {
    __amdgpu_buffer_rsrc_t a, lo, off;
    uint32_t a_lane, lo_lane;
    __device__ __forceinline__ void load(int it, RingItem& r) const
    {
        const auto av = __builtin_amdgcn_raw_buffer_load_b64(a, a_lane, 512 * it, 0);
        r.ax = __uint_as_float(av[0]);
        r.ay = __uint_as_float(av[1]);
        const auto lv = __builtin_amdgcn_raw_buffer_load_b64(lo, lo_lane, 256 * it, 0);
        r.lo = make_uint2(lv[0], lv[1]);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const auto ov = __builtin_amdgcn_raw_buffer_load_b128(off, 0u, 64 * it + 16 * q, 0);
            r.o[4 * q + 0] = ov[0];
            r.o[4 * q + 1] = ov[1];
            r.o[4 * q + 2] = ov[2];
            r.o[4 * q + 3] = ov[3];
        }
    }
};

// the 4 products of pair group g (pairs 4g .. 4g + 3) into the group's own
// sum: ABID 4 (g & 1) + q of the couplings, ABID g of the low parts
// step q of the four pair groups of a full item, one MFMA per group's sum
template <int Q>
__device__ __forceinline__ void ring_step(const RingItem& it, const float* x, rf4* acc)
{
    acc[0] = __builtin_amdgcn_mfma_f32_4x4x1f32(it.ax, x[Q], acc[0], 3, Q, 0);
    acc[1] = __builtin_amdgcn_mfma_f32_4x4x1f32(it.ax, x[4 + Q], acc[1], 3, 4 + Q, 0);
    acc[2] = __builtin_amdgcn_mfma_f32_4x4x1f32(it.ay, x[8 + Q], acc[2], 3, Q, 0);
    acc[3] = __builtin_amdgcn_mfma_f32_4x4x1f32(it.ay, x[12 + Q], acc[3], 3, 4 + Q, 0);
}

template <int G>
__device__ __forceinline__ void ring_lo(const RingItem& it, const float* x, rf4& acl)
{
    const uint2 xb = {ring_bf16_pair(x[4 * G], x[4 * G + 1]), ring_bf16_pair(x[4 * G + 2], x[4 * G + 3])};
    acl = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(__builtin_bit_cast(rs4, it.lo), __builtin_bit_cast(rs4, xb), acl,
                                                 3, G, 0);
}

template <int G>
__device__ __forceinline__ void ring_group(const RingItem& it, const float* x, rf4& acc, rf4& acl)
{
    const float av = G < 2 ? it.ax : it.ay;
    constexpr int B = 4 * (G & 1);
    acc = __builtin_amdgcn_mfma_f32_4x4x1f32(av, x[4 * G + 0], acc, 3, B + 0, 0);
    acc = __builtin_amdgcn_mfma_f32_4x4x1f32(av, x[4 * G + 1], acc, 3, B + 1, 0);
    acc = __builtin_amdgcn_mfma_f32_4x4x1f32(av, x[4 * G + 2], acc, 3, B + 2, 0);
    acc = __builtin_amdgcn_mfma_f32_4x4x1f32(av, x[4 * G + 3], acc, 3, B + 3, 0);
    const uint2 xb = {ring_bf16_pair(x[4 * G], x[4 * G + 1]), ring_bf16_pair(x[4 * G + 2], x[4 * G + 3])};
    acl = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(__builtin_bit_cast(rs4, it.lo), __builtin_bit_cast(rs4, xb), acl,
                                                 3, G, 0);
}

template <bool EXACT>
__global__ __launch_bounds__(1024) void k_spmm_ring(Dev d, int j, int last)
{
    extern __shared__ float ring[];                     // [32 columns][kRingStride], then red
    double* red = reinterpret_cast<double*>(ring + 32 * kRingStride + (32 * kRingStride & 1));
    const int nsl = d.sn / kRingTS;
    const int logical = xcd_tile(blockIdx.x, gridDim.x);
    if (logical >= d.ntile_p * nsl) return;
    // range-major: the slices of one row range run on one XCD and share its
    // couplings through that XCD's L2
    const int range = logical / nsl;
    const int bs = d.s0 + (logical % nsl) * kRingTS;    // first trajectory slot of the slice
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = uniform(tid >> 6), nw = (int)(blockDim.x >> 6);
    const int h = lane >> 5, cl = lane & 31;            // pair half, real column in the slice
    const int Bp = d.Bp, n = d.n;
    const int slice_nact = uniform(d.nact[bs]);         // slots are sorted by n_b: the slice's largest
    const int lim = min(n, slice_nact);                 // rows holding data in this slice
    const int R0 = range * d.ring_rows;
    const int R1 = min(min(n, R0 + d.ring_rows), slice_nact);
    // x, w and the diagonal of this slice as buffers that end with row n - 1
    // (the groups of 4 rows past it read 0): row r at byte r * 8 Bp (4 Bp),
    // this lane's column at 4 cl (n * 8 Bp < 2^31, checked on the host)
    const uint32_t rowb = 8u * (uint32_t)Bp;
    const int vbytes = (int)((uint32_t)n * rowb - 8u * (uint32_t)bs);
    const __amdgpu_buffer_rsrc_t xr =
        ring_rsrc(reinterpret_cast<const float2*>(d.V) + (size_t)j * d.vs + bs, vbytes);
    const __amdgpu_buffer_rsrc_t yr =
        ring_rsrc(reinterpret_cast<const float2*>(d.V) + (size_t)(j + 1) * d.vs + bs, vbytes);
    const __amdgpu_buffer_rsrc_t dr =
        ring_rsrc(reinterpret_cast<const float*>(d.diag) + bs, (int)((uint32_t)n * 4u * (uint32_t)Bp - 4u * (uint32_t)bs));
    const uint32_t col_lds = (uint32_t)cl * (kRingStride * 4);
    double a0 = 0.0, a1 = 0.0;                           // FP64 partial <x|w> (re: a0, im: a1 / |w|^2: a0)
    if (R0 < R1) {
        // a group of 4 rows from r: lane (h, cl) holds rows r + 2h, r + 2h + 1 of column cl
        const uint32_t xlane = 4u * (uint32_t)cl + 2u * (uint32_t)h * rowb;
        auto load_grp = [&](int r, float& v0, float& v1) {
            const int ra = r + 2 * h;
            const uint32_t so = (uint32_t)r * rowb;
            v0 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(xr, xlane, so, 0));
            v1 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(xr, xlane + rowb, so, 0));
            if (ra >= lim) v0 = 0.f;
            if (ra + 1 >= lim) v1 = 0.f;
        };
        auto store_grp = [&](int r, float v0, float v1) {
            const int ra = (r + 2 * h) & (kRingCap - 1);
            ring[cl * kRingStride + ra] = v0;
            ring[cl * kRingStride + ra + 1] = v1;
        };
        const int c0 = R0 / kRingChunk, c1 = (R1 + kRingChunk - 1) / kRingChunk, nc = c1 - c0;
        const int k0 = R0 >> 2, nbk = ((R1 + 3) >> 2) - k0;
        // the range's chunk windows and block item indices, once, in LDS
        int2* twin = reinterpret_cast<int2*>(red + nw * 32);
        int* titem = reinterpret_cast<int*>(twin + nc);
        for (int i = tid; i < nc; i += (int)blockDim.x) twin[i] = d.ring_win[c0 + i];
        for (int i = tid; i <= nbk; i += (int)blockDim.x) titem[i] = d.ring_item[k0 + i];
        __syncthreads();
        // a chunk's window, 4 groups of 4 rows in flight per wave
        auto prefill = [&](int c) {
            const int2 w = twin[c - c0];
            const int G = (w.y - w.x) >> 2;
            for (int g = wave; g < G; g += 4 * nw) {
                float v[4][2];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    v[u][0] = v[u][1] = 0.f;
                    if (g + u * nw < G) load_grp(w.x + 4 * (g + u * nw), v[u][0], v[u][1]);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (g + u * nw < G) store_grp(w.x + 4 * (g + u * nw), v[u][0], v[u][1]);
            }
        };
        const double gc = d.g;
        const int tl = cl >> 1, comp = cl & 1;           // trajectory in the slice, re/im
        const int na = d.nact[bs + tl];
        const uint32_t lb = col_lds + 4u * (uint32_t)h;
        const int bpw = (kRingChunk / 4) / nw;           // consecutive blocks per wave and chunk
        RingSrc src;
        src.a = ring_rsrc(d.ring_a);
        src.lo = ring_rsrc(d.ring_alo);
        src.off = ring_rsrc(d.ring_off);
        src.a_lane = 8u * (uint32_t)lane;
        src.lo_lane = 8u * (uint32_t)((((lane >> 2) & 3) * 2 + h) * 4 + (lane & 3));
        // a block's diagonal (rows 4k + 2h, + 1 for this lane's trajectory)
        const uint32_t dlane = 4u * (uint32_t)tl + 2u * (uint32_t)h * (4u * (uint32_t)Bp);
        float dv0 = 0.f, dv1 = 0.f;
        auto load_diag = [&](int kk) {
            const uint32_t so = 4u * (uint32_t)kk * (4u * (uint32_t)Bp);
            dv0 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(dr, dlane, so, 0));
            dv1 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(dr, dlane + 4u * (uint32_t)Bp, so, 0));
        };
        // one FP32 sum per pair group (independent chains on the matrix
        // cores), the low parts' sum, FP64 sums per block
        rf4 acc[4], acl = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int g = 0; g < 4; ++g) acc[g] = rf4{0.f, 0.f, 0.f, 0.f};
        double s[4] = {0.0, 0.0, 0.0, 0.0};
        // this wave's blocks [k, kB) of the chunk, and of the next chunk of the
        // segment (nkA; nxt_has: it has some)
        int k = 0, kB = 0, nkA = 0;
        bool nxt_has = false;
        // the x values of an item's pairs (LDS reads issued one item ahead)
        auto read_x = [&](const RingItem& it, float* x) {
            const uint32_t meta = (uint32_t)uniform((int)it.o[0]) & 7u;
            const int np = 4 * (int)((meta & 3u) + 1u);
            x[0] = lds_f32(ring, lb + (it.o[0] & ~7u));
#pragma unroll
            for (int q = 1; q < 4; ++q) x[q] = lds_f32(ring, lb + it.o[q]);
#pragma unroll
            for (int g = 1; g < 4; ++g)
                if (np > 4 * g) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) x[4 * g + q] = lds_f32(ring, lb + it.o[4 * g + q]);
                }
        };
        // the item's products, the next item's x reads (not across a chunk
        // boundary: the ring moves there), then the FP64 sums and, at a
        // block's last item, its rows
        auto process = [&](const RingItem& cur, const float* x, const RingItem& nxt, float* xn, bool read_next) {
            const uint32_t meta = (uint32_t)uniform((int)cur.o[0]) & 7u;
            const int np = 4 * (int)((meta & 3u) + 1u);     // pairs in the item: 4 .. 16
            if (np == 16) {
                // a full item: the four groups' chains interleaved
                ring_step<0>(cur, x, acc);
                ring_step<1>(cur, x, acc);
                ring_step<2>(cur, x, acc);
                ring_step<3>(cur, x, acc);
                ring_lo<0>(cur, x, acl);
                ring_lo<1>(cur, x, acl);
                ring_lo<2>(cur, x, acl);
                ring_lo<3>(cur, x, acl);
            } else {
                ring_group<0>(cur, x, acc[0], acl);
                if (np > 4) ring_group<1>(cur, x, acc[1], acl);
                if (np > 8) ring_group<2>(cur, x, acc[2], acl);
            }
            if (read_next) read_x(nxt, xn);
            if (EXACT) {
                // FP64 sums of every 4 products
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    if (g > 0 && np <= 4 * g) break;
#pragma unroll
                    for (int i = 0; i < 4; ++i) s[i] += (double)acc[g][i];
                }
            } else {
                // the groups' FP32 sums added in FP32, then into FP64
#pragma unroll
                for (int i = 0; i < 4; ++i) s[i] += (double)((acc[0][i] + acc[1][i]) + (acc[2][i] + acc[3][i]));
            }
#pragma unroll
            for (int g = 0; g < 4; ++g) acc[g] = rf4{0.f, 0.f, 0.f, 0.f};
            if (meta & 4u) {
#pragma unroll
                for (int i = 0; i < 4; ++i) s[i] += (double)acl[i];
                // half h finishes rows 2h, 2h + 1: its own sums plus the other half's
                const double send0 = h ? s[0] : s[2], send1 = h ? s[1] : s[3];
                const double got0 = __shfl_xor(send0, 32, 64), got1 = __shfl_xor(send1, 32, 64);
                const double t0 = (h ? s[2] : s[0]) + got0, t1 = (h ? s[3] : s[1]) + got1;
                const int r = 4 * k + 2 * h;
                const uint32_t xo = col_lds + 4u * (uint32_t)(r & (kRingCap - 1));
                const float x0 = lds_f32(ring, xo), x1 = lds_f32(ring, xo + 4u);
                float w0 = 0.f, w1 = 0.f;
                if (r < n && r < na) w0 = (float)__builtin_fma(gc, t0, (double)dv0 * (double)x0);
                if (r + 1 < n && r + 1 < na) w1 = (float)__builtin_fma(gc, t1, (double)dv1 * (double)x1);
                const double dx0 = x0, dx1 = x1, dw0 = w0, dw1 = w1;
                if (!last) {
                    const uint32_t so = 4u * (uint32_t)k * rowb;
                    if (r < n) __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(w0), yr, xlane, so, 0);
                    if (r + 1 < n) __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(w1), yr, xlane + rowb, so, 0);
                    // the other component of the same trajectory is in lane ^ 1
                    const double wp0 = swap_pair(w0), wp1 = swap_pair(w1);
                    const double sx0 = comp ? -dx0 : dx0, sx1 = comp ? -dx1 : dx1;
                    a0 = __builtin_fma(dx1, dw1, __builtin_fma(dx0, dw0, a0));
                    a1 = __builtin_fma(sx1, wp1, __builtin_fma(sx0, wp0, a1));
                } else {
                    a0 = __builtin_fma(dw1, dw1, __builtin_fma(dw0, dw0, a0));
                }
                ++k;
                // the next block's diagonal (the next chunk's first, after the last)
                if (k < kB) load_diag(k);
                else if (nxt_has) load_diag(nkA);
                acl = rf4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int i = 0; i < 4; ++i) s[i] = 0.0;
            }
        };
        RingItem b0, b1;
        float x0[16], x1[16];
        // chunks [ca, cb) in order: the window of ca first, then each chunk's
        // new rows loaded while the one before multiplies
        auto segment = [&](int ca, int cb) {
            prefill(ca);
            __syncthreads();
            bool have = false;                           // b0 and dv0/dv1 hold this chunk's first
            for (int c = ca; c < cb; ++c) {
                // rows the next chunk adds, [hi_c, hi_{c+1}): a wave's first two
                // groups are loaded now and stored after this chunk's products
                // (the ring slots they take hold no row of this chunk's window)
                int gbase = 0, ng = 0;
                float p0 = 0.f, p1 = 0.f, q0 = 0.f, q1 = 0.f;
                if (c + 1 < cb) {
                    gbase = twin[c - c0].y;
                    ng = (twin[c + 1 - c0].y - gbase) >> 2;
                    if (wave < ng) load_grp(gbase + 4 * wave, p0, p1);
                    if (wave + nw < ng) load_grp(gbase + 4 * (wave + nw), q0, q1);
                }
                const int kA = c * (kRingChunk / 4) + wave * bpw;
                kB = min(kA + bpw, (min(R1, (c + 1) * kRingChunk) + 3) >> 2);
                nkA = (c + 1) * (kRingChunk / 4) + wave * bpw;
                nxt_has = c + 1 < cb && nkA < min(nkA + bpw, (min(R1, (c + 2) * kRingChunk) + 3) >> 2);
                if (kA < kB) {
                    int it = titem[kA - k0];
                    const int it_end = titem[kB - k0];
                    const int nfirst = nxt_has ? titem[nkA - k0] : it_end - 1;
                    k = kA;
                    if (!have) {
                        src.load(it, b0);
                        load_diag(k);
                    }
                    read_x(b0, x0);
                    // two item buffers in turn: the next item loads while this
                    // one multiplies, and its x reads follow this one's products
                    for (;;) {
                        src.load(it + 1 < it_end ? it + 1 : nfirst, b1);
                        process(b0, x0, b1, x1, it + 1 < it_end);
                        if (++it == it_end) {
                            b0 = b1;                     // the next chunk's first item
                            break;
                        }
                        src.load(it + 1 < it_end ? it + 1 : nfirst, b0);
                        process(b1, x1, b0, x0, it + 1 < it_end);
                        if (++it == it_end) break;
                    }
                    have = nxt_has;
                } else {
                    have = false;
                }
                if (c + 1 < cb) {
                    if (wave < ng) store_grp(gbase + 4 * wave, p0, p1);
                    if (wave + nw < ng) store_grp(gbase + 4 * (wave + nw), q0, q1);
                    for (int g = wave + 2 * nw; g < ng; g += nw) {
                        float v0, v1;
                        load_grp(gbase + 4 * g, v0, v1);
                        store_grp(gbase + 4 * g, v0, v1);
                    }
                }
                __syncthreads();
            }
        };
        // the slices of a range start at different chunks, so that the slices
        // sharing an XCD do not all read a chunk's couplings at the same time
        // (a slot's slice number, not its place in the launch: results do not
        // depend on how the batch is split into launches)
        const int rot = (bs / kRingTS) % nc;
        segment(c0 + rot, c1);
        if (rot > 0) segment(c0, c0 + rot);
    }
    // per trajectory: the re/im lanes, the pair halves, then the waves in order
    a0 += __shfl_xor(a0, 1, 64);
    a1 += __shfl_xor(a1, 1, 64);
    a0 += __shfl_xor(a0, 32, 64);
    a1 += __shfl_xor(a1, 32, 64);
    if (lane < 32 && (lane & 1) == 0) {
        red[(wave * 2 + 0) * 16 + (lane >> 1)] = a0;
        red[(wave * 2 + 1) * 16 + (lane >> 1)] = a1;
    }
    __syncthreads();
    if (wave == 0 && lane < 16) {
        for (int q = 0; q < 2; ++q) {
            if (q == 1 && last) break;
            double v = 0.0;
            for (int w = 0; w < nw; ++w) v += red[(w * 2 + q) * 16 + lane];
            d.part_s[((size_t)q * d.ntile_p + range) * Bp + bs + lane] = v;
        }
    }
}

static inline int rpad8(int x) { return (x + 7) & ~7; }

hipError_t launch_spmm_ring(const Dev& d, int j, int last, hipStream_t s)
{
    const dim3 grid(rpad8(d.ntile_p * (d.sn / kRingTS))), block(64 * d.ring_w);
    const size_t lds = spmm_ring_lds_bytes(d.ring_w, d.ring_rows);
    if (d.spmm_exact) hipLaunchKernelGGL((k_spmm_ring<true>), grid, block, lds, s, d, j, last);
    else hipLaunchKernelGGL((k_spmm_ring<false>), grid, block, lds, s, d, j, last);
    return hipGetLastError();
}

// lift the dynamic LDS limit of both instantiations (once per process)
hipError_t spmm_ring_set_lds_limit()
{
    const int bytes = (int)spmm_ring_lds_bytes(16, kRingMaxRows);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_spmm_ring<false>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e == hipSuccess)
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_spmm_ring<true>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    return e;
}

}  // namespace dq
