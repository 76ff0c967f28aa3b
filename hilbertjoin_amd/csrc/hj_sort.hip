// hj_sort.hip -- device exclusive scans and the LSD radix sort of 64-bit keys.
//
// The radix sort replaces the reference's EGO quicksort of whole FP64 rows
// (omp_qsort egojoin.cpp:594-634 driven by epsilonGridOrdering :159-168):
// instead of moving d*8-byte rows through n log n comparisons, the grid cells
// are packed into one 64-bit key per point and (key, index) pairs are sorted
// with 8-bit LSD passes; the rows are gathered once afterwards.
//
// One pass = three launches:
//   upsweep   per-block digit histogram          hist[digit * nblocks + block]
//   row scan  one block per digit: exclusive scan of its row over the blocks,
//             and the digit's total
//   downsweep digit bases (exclusive scan of the 256 totals, per block in
//             LDS), stable local ranking (ballot-based peer matching in each
//             wavefront), reorder of the block's tile in LDS by digit, then a
//             coalesced write-out: consecutive lanes write consecutive slots of
//             a digit run.
#include "hj_internal.h"

namespace hj {

namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;  // 2048

template <class T>
__device__ T block_exclusive_scan(T v, T *sh, T *total) {
    // Hillis-Steele over 256 threads in LDS (sh has 256 entries)
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_THREADS; off <<= 1) {
        T add = t >= off ? sh[t - off] : T(0);
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    T incl = sh[t];
    if (total) *total = sh[SCAN_THREADS - 1];
    __syncthreads();
    return incl - v;
}

template <class T>
__global__ __launch_bounds__(SCAN_THREADS) void scan_reduce_kernel(const T *in, size_t n,
                                                                   T *partials) {
    __shared__ T sh[SCAN_THREADS];
    size_t base = (size_t)blockIdx.x * SCAN_TILE;
    T s = 0;
    for (int i = 0; i < SCAN_ITEMS; i++) {
        size_t idx = base + (size_t)i * SCAN_THREADS + threadIdx.x;
        if (idx < n) s += in[idx];
    }
    T tot;
    block_exclusive_scan<T>(s, sh, &tot);
    if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

template <class T>
__global__ __launch_bounds__(SCAN_THREADS) void scan_apply_kernel(const T *in, T *out, size_t n,
                                                                  const T *partials_scanned) {
    __shared__ T sh[SCAN_THREADS];
    size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
    T v[SCAN_ITEMS];
    T s = 0;
    for (int i = 0; i < SCAN_ITEMS; i++) {
        size_t idx = base + i;
        v[i] = idx < n ? in[idx] : T(0);
        s += v[i];
    }
    T ex = block_exclusive_scan<T>(s, sh, nullptr);
    T run = partials_scanned[blockIdx.x] + ex;
    for (int i = 0; i < SCAN_ITEMS; i++) {
        size_t idx = base + i;
        if (idx < n) out[idx] = run;
        run += v[i];
    }
}

// Up to SCAN_DIRECT_BLOCKS blocks the partials need no scan of their own:
// the apply kernel's block b sums partials[0, b) itself (one per thread,
// then a block reduction), so a scan takes two launches.
constexpr size_t SCAN_DIRECT_BLOCKS = SCAN_THREADS * 4;

template <class T>
__global__ __launch_bounds__(SCAN_THREADS) void scan_apply_direct_kernel(const T *in, T *out,
                                                                         size_t n,
                                                                         const T *partials) {
    __shared__ T sh[SCAN_THREADS];
    T before = 0;
    for (uint32_t q = threadIdx.x; q < blockIdx.x; q += SCAN_THREADS) before += partials[q];
    T tot;
    block_exclusive_scan<T>(before, sh, &tot);  // tot: the sum over all threads
    size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
    T v[SCAN_ITEMS];
    T s = 0;
    for (int i = 0; i < SCAN_ITEMS; i++) {
        size_t idx = base + i;
        v[i] = idx < n ? in[idx] : T(0);
        s += v[i];
    }
    T ex = block_exclusive_scan<T>(s, sh, nullptr);
    T run = tot + ex;
    for (int i = 0; i < SCAN_ITEMS; i++) {
        size_t idx = base + i;
        if (idx < n) out[idx] = run;
        run += v[i];
    }
}

template <class T>
void exclusive_scan_impl(const T *in, T *out, size_t n, DevBuf &tmp, size_t tmp_off,
                         hipStream_t s) {
    if (n == 0) return;
    size_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    // partials for this level live at tmp[tmp_off .. tmp_off + 2*nb)
    size_t need = (tmp_off + 2 * nb + 64) * sizeof(T);
    if (tmp.bytes < need) throw InvalidArg{"scan scratch too small"};
    T *part = tmp.as<T>() + tmp_off;
    T *part_scanned = part + nb;
    if (nb > 1) {
        scan_reduce_kernel<T><<<(unsigned)nb, SCAN_THREADS, 0, s>>>(in, n, part);
        HJ_HIP(hipGetLastError());
    }
    if (nb <= SCAN_DIRECT_BLOCKS) {
        scan_apply_direct_kernel<T><<<(unsigned)nb, SCAN_THREADS, 0, s>>>(in, out, n, part);
        HJ_HIP(hipGetLastError());
        return;
    }
    exclusive_scan_impl<T>(part, part_scanned, nb, tmp, tmp_off + 2 * nb, s);
    scan_apply_kernel<T><<<(unsigned)nb, SCAN_THREADS, 0, s>>>(in, out, n, part_scanned);
    HJ_HIP(hipGetLastError());
}

template <class T> size_t scan_tmp_elems(size_t n) {
    size_t tot = 64;
    while (n > 1) {
        size_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
        tot += 2 * nb;
        n = nb;
    }
    return tot + 64;
}

// ---------------------------------------------------------------------------
// radix sort
// ---------------------------------------------------------------------------

constexpr int RS_THREADS = 256;
constexpr int RS_WAVES = RS_THREADS / 64;
constexpr int RS_RADIX = 256;
// keys per thread: 16 (4096 keys per block) for large sorts, 4 (1024 keys per
// block) below RS_SMALL_N keys so that a sort of about a million keys still
// spreads over a thousand blocks
constexpr int RS_ITEMS_LARGE = 16;
constexpr int RS_ITEMS_SMALL = 4;
constexpr size_t RS_SMALL_N = (size_t)4 << 20;

template <int ITEMS>
__global__ __launch_bounds__(RS_THREADS) void rs_upsweep(const uint64_t *__restrict__ keys,
                                                         size_t n, int shift,
                                                         uint32_t *__restrict__ hist,
                                                         uint32_t nblocks) {
    __shared__ uint32_t h[RS_WAVES][RS_RADIX];
    const int t = threadIdx.x, w = t >> 6;
    for (int i = t; i < RS_WAVES * RS_RADIX; i += RS_THREADS) (&h[0][0])[i] = 0;
    __syncthreads();
    constexpr int TILE = RS_THREADS * ITEMS;
    size_t base = (size_t)blockIdx.x * TILE;
    for (int it = 0; it < ITEMS; it++) {
        size_t idx = base + (size_t)it * RS_THREADS + t;
        if (idx < n) atomicAdd(&h[w][(uint32_t)(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    uint32_t c = 0;
    for (int ww = 0; ww < RS_WAVES; ww++) c += h[ww][t];
    hist[(size_t)t * nblocks + blockIdx.x] = c;
}

// One block per digit: exclusive scan of hist[digit][0..nblocks) into
// offs[digit][..] and the digit's total into totals[digit].
__global__ __launch_bounds__(RS_THREADS) void rs_scan_rows(const uint32_t *__restrict__ hist,
                                                           uint32_t nblocks,
                                                           uint32_t *__restrict__ offs,
                                                           uint32_t *__restrict__ totals) {
    __shared__ uint32_t sh[RS_THREADS];
    const uint32_t dg = blockIdx.x;
    const uint32_t *row = hist + (size_t)dg * nblocks;
    uint32_t *out = offs + (size_t)dg * nblocks;
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nblocks; b0 += RS_THREADS) {
        const uint32_t b = b0 + threadIdx.x;
        const uint32_t v = b < nblocks ? row[b] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan<uint32_t>(v, sh, &tot);
        if (b < nblocks) out[b] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) totals[dg] = carry;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    const uint32_t lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

template <class V, int ITEMS>
__global__ __launch_bounds__(RS_THREADS) void rs_downsweep(
    const uint64_t *__restrict__ kin, const V *__restrict__ vin, uint64_t *__restrict__ kout,
    V *__restrict__ vout, size_t n, int shift, const uint32_t *__restrict__ offs,
    const uint32_t *__restrict__ totals, uint32_t nblocks) {
    __shared__ uint32_t wave_cnt[RS_WAVES][RS_RADIX];
    __shared__ uint32_t dstart[RS_RADIX];
    __shared__ uint32_t goff[RS_RADIX];
    __shared__ uint32_t scan_sh[RS_THREADS];
    constexpr int TILE = RS_THREADS * ITEMS;
    __shared__ uint64_t skeys[TILE];
    __shared__ V svals[TILE];

    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    for (int i = t; i < RS_WAVES * RS_RADIX; i += RS_THREADS) (&wave_cnt[0][0])[i] = 0;
    __syncthreads();

    uint64_t k[ITEMS];
    V v[ITEMS];
    uint32_t rank[ITEMS];
    const size_t base = (size_t)blockIdx.x * TILE + (size_t)w * (ITEMS * 64);
    const uint64_t lt = lanemask_lt();
    for (int it = 0; it < ITEMS; it++) {
        size_t idx = base + (size_t)it * 64 + lane;
        bool valid = idx < n;
        k[it] = valid ? kin[idx] : 0ull;
        v[it] = valid ? vin[idx] : V(0);
        uint32_t dg = (uint32_t)(k[it] >> shift) & 255u;
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < 8; b++) {
            bool bit = (dg >> b) & 1u;
            uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        uint32_t before = (uint32_t)__popcll(peers & lt);
        uint32_t cnt = (uint32_t)__popcll(peers);
        uint32_t basec = valid ? wave_cnt[w][dg] : 0u;
        rank[it] = basec + before;
        // all peers have read the counter above; the lowest peer advances it
        __builtin_amdgcn_wave_barrier();
        if (valid && before == 0) wave_cnt[w][dg] = basec + cnt;
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // per digit: exclusive over waves, block total, block-local digit start
    uint32_t tot = 0;
    for (int ww = 0; ww < RS_WAVES; ww++) {
        uint32_t c = wave_cnt[ww][t];
        wave_cnt[ww][t] = tot;
        tot += c;
    }
    uint32_t dtot;
    // global base of digit t: the totals of all smaller digits
    const uint32_t gbase = block_exclusive_scan<uint32_t>(totals[t], scan_sh, &dtot);
    goff[t] = gbase + offs[(size_t)t * nblocks + blockIdx.x];
    uint32_t ex = block_exclusive_scan<uint32_t>(tot, scan_sh, &dtot);
    dstart[t] = ex;
    __syncthreads();
    for (int it = 0; it < ITEMS; it++) {
        size_t idx = base + (size_t)it * 64 + lane;
        if (idx < n) {
            uint32_t dg = (uint32_t)(k[it] >> shift) & 255u;
            uint32_t pos = dstart[dg] + wave_cnt[w][dg] + rank[it];
            skeys[pos] = k[it];
            svals[pos] = v[it];
        }
    }
    __syncthreads();
    size_t start = (size_t)blockIdx.x * TILE;
    uint32_t nvalid = (uint32_t)((n - start) < (size_t)TILE ? (n - start) : TILE);
    for (uint32_t p = t; p < nvalid; p += RS_THREADS) {
        uint64_t key = skeys[p];
        uint32_t dg = (uint32_t)(key >> shift) & 255u;
        uint32_t o = goff[dg] + (p - dstart[dg]);
        kout[o] = key;
        vout[o] = svals[p];
    }
}

template <class V, int ITEMS>
int radix_passes(uint64_t *keys[2], V *vals[2], size_t n, int bits, DevBuf &hist,
                 hipStream_t s) {
    constexpr size_t TILE = (size_t)RS_THREADS * ITEMS;
    const uint32_t nblocks = (uint32_t)((n + TILE - 1) / TILE);
    const size_t hn = (size_t)RS_RADIX * nblocks;
    hist.ensure((2 * hn + RS_RADIX) * sizeof(uint32_t));
    uint32_t *h = hist.as<uint32_t>();
    uint32_t *o = h + hn;
    uint32_t *tot = o + hn;
    int cur = 0;
    for (int shift = 0; shift < bits; shift += 8) {
        rs_upsweep<ITEMS><<<nblocks, RS_THREADS, 0, s>>>(keys[cur], n, shift, h, nblocks);
        HJ_HIP(hipGetLastError());
        rs_scan_rows<<<RS_RADIX, RS_THREADS, 0, s>>>(h, nblocks, o, tot);
        HJ_HIP(hipGetLastError());
        rs_downsweep<V, ITEMS><<<nblocks, RS_THREADS, 0, s>>>(
            keys[cur], vals[cur], keys[cur ^ 1], vals[cur ^ 1], n, shift, o, tot, nblocks);
        HJ_HIP(hipGetLastError());
        cur ^= 1;
    }
    return cur;
}

template <class V>
void radix_sort_impl(uint64_t *keys[2], V *vals[2], size_t n, int bits, int *which,
                     DevBuf &hist, DevBuf &scan_tmp, hipStream_t s) {
    (void)scan_tmp;
    int cur = 0;
    if (n > 1 && bits > 0) {
        if (n < RS_SMALL_N)
            cur = radix_passes<V, RS_ITEMS_SMALL>(keys, vals, n, bits, hist, s);
        else
            cur = radix_passes<V, RS_ITEMS_LARGE>(keys, vals, n, bits, hist, s);
    }
    *which = cur;
}

}  // namespace

void exclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n, DevBuf &tmp,
                        hipStream_t s) {
    tmp.ensure(scan_tmp_elems<uint32_t>(n) * sizeof(uint32_t));
    exclusive_scan_impl<uint32_t>(in, out, n, tmp, 0, s);
}

void exclusive_scan_u64(const uint64_t *in, uint64_t *out, size_t n, DevBuf &tmp,
                        hipStream_t s) {
    tmp.ensure(scan_tmp_elems<uint64_t>(n) * sizeof(uint64_t));
    exclusive_scan_impl<uint64_t>(in, out, n, tmp, 0, s);
}

void radix_sort_u64_u32(uint64_t *keys[2], uint32_t *vals[2], size_t n, int bits, int *which,
                        DevBuf &hist, DevBuf &scan_tmp, hipStream_t s) {
    radix_sort_impl<uint32_t>(keys, vals, n, bits, which, hist, scan_tmp, s);
}

void radix_sort_u64_u64(uint64_t *keys[2], uint64_t *vals[2], size_t n, int bits, int *which,
                        DevBuf &hist, DevBuf &scan_tmp, hipStream_t s) {
    radix_sort_impl<uint64_t>(keys, vals, n, bits, which, hist, scan_tmp, s);
}

}  // namespace hj
