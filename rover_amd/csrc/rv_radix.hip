// rv_radix.hip -- stable LSD radix sort of 64-bit keys (optionally with
// 32-bit values) for the graph builders: 8-bit digits; per pass a count
// kernel (digit histogram of each 4096-key tile in LDS), an exclusive scan
// of the digit-major counts, and a scatter that keeps the input order among
// equal digits (tiles in order, 256 keys per round in thread order; the
// rank among equal digits inside a wavefront comes from eight ballots, the
// wavefronts before it from LDS counts).  Keys and values ping-pong between
// the caller's output arrays and the scratch.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "rover_hip.h"

#define RX_THREADS 256
#define RX_WAVES (RX_THREADS / 64)
#define RX_ITEMS 16
#define RX_TILE (RX_THREADS * RX_ITEMS)
#define RX_BINS 256

static int rx_check(hipError_t e, const char *what) {
    if (e == hipSuccess) return 0;
    char msg[256];
    snprintf(msg, sizeof(msg), "%s: %s", what, hipGetErrorString(e));
    rv_set_error(msg);
    return -1;
}

__device__ __forceinline__ int rx_digit(uint64_t k, int shift, uint32_t mask) {
    return (int)((k >> shift) & mask);
}

// counts[d * tiles + tile] = keys of digit d in the tile
__global__ __launch_bounds__(RX_THREADS) void rx_count_kernel(const uint64_t *__restrict__ keys,
                                                              size_t n, int shift, uint32_t mask,
                                                              uint32_t *__restrict__ counts,
                                                              int tiles) {
    __shared__ uint32_t h[RX_BINS];
    const int t = threadIdx.x;
    h[t] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * RX_TILE;
    for (int i = 0; i < RX_ITEMS; i++) {
        const size_t idx = base + (size_t)i * RX_THREADS + t;
        if (idx < n) atomicAdd(&h[rx_digit(keys[idx], shift, mask)], 1u);
    }
    __syncthreads();
    counts[(size_t)t * tiles + blockIdx.x] = h[t];
}

// Exclusive scan of one 4096-entry block of v in place; the block's total
// goes to sums[blockIdx.x] when sums is given.
__global__ __launch_bounds__(RX_THREADS) void rx_scan_block_kernel(uint32_t *__restrict__ v,
                                                                   size_t m,
                                                                   uint32_t *__restrict__ sums) {
    __shared__ uint32_t part[RX_THREADS];
    const int t = threadIdx.x;
    const size_t base = (size_t)blockIdx.x * RX_TILE + (size_t)t * RX_ITEMS;
    uint32_t local[RX_ITEMS];
    uint32_t s = 0;
    for (int i = 0; i < RX_ITEMS; i++) {
        const size_t idx = base + i;
        local[i] = idx < m ? v[idx] : 0u;
        s += local[i];
    }
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < RX_THREADS; off <<= 1) {
        const uint32_t add = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    uint32_t run = part[t] - s;
    for (int i = 0; i < RX_ITEMS; i++) {
        const size_t idx = base + i;
        if (idx < m) v[idx] = run;
        run += local[i];
    }
    if (sums && t == RX_THREADS - 1) sums[blockIdx.x] = part[t];
}

// v[i] += offs[i / 4096]
__global__ void rx_add_kernel(uint32_t *__restrict__ v, size_t m,
                              const uint32_t *__restrict__ offs) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) v[i] += offs[i / RX_TILE];
}

// Scratch words the exclusive scan of m entries needs for its block sums.
static size_t rx_scan_words(size_t m) {
    size_t w = 0;
    while (m > RX_TILE) {
        m = (m + RX_TILE - 1) / RX_TILE;
        w += m;
    }
    return w;
}

static int rx_scan(uint32_t *v, size_t m, uint32_t *work, hipStream_t s) {
    if (m <= RX_TILE) {
        rx_scan_block_kernel<<<1, RX_THREADS, 0, s>>>(v, m, nullptr);
        return rx_check(hipGetLastError(), "rx_scan_block_kernel");
    }
    const size_t blocks = (m + RX_TILE - 1) / RX_TILE;
    uint32_t *sums = work;
    rx_scan_block_kernel<<<(unsigned)blocks, RX_THREADS, 0, s>>>(v, m, sums);
    if (rx_check(hipGetLastError(), "rx_scan_block_kernel")) return -1;
    if (rx_scan(sums, blocks, work + blocks, s)) return -1;
    rx_add_kernel<<<(unsigned)((m + 255) / 256), 256, 0, s>>>(v, m, sums);
    return rx_check(hipGetLastError(), "rx_add_kernel");
}

__global__ __launch_bounds__(RX_THREADS) void rx_scatter_kernel(
    const uint64_t *__restrict__ kin, uint64_t *__restrict__ kout,
    const uint32_t *__restrict__ vin, uint32_t *__restrict__ vout, size_t n, int shift,
    uint32_t mask, const uint32_t *__restrict__ offs, int tiles) {
    __shared__ uint32_t next[RX_BINS];              // this tile's next slot per digit
    __shared__ uint32_t wcnt[RX_WAVES][RX_BINS];    // keys per digit and wave, this round
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    next[t] = offs[(size_t)t * tiles + blockIdx.x];
    const uint64_t below = (1ull << lane) - 1ull;
    const size_t base = (size_t)blockIdx.x * RX_TILE;
    for (int r = 0; r < RX_ITEMS; r++) {
        for (int w = 0; w < RX_WAVES; w++) wcnt[w][t] = 0u;
        __syncthreads();
        const size_t idx = base + (size_t)r * RX_THREADS + t;
        const bool valid = idx < n;
        const uint64_t key = valid ? kin[idx] : 0ull;
        const int d = valid ? rx_digit(key, shift, mask) : 0;
        // lanes of this wave holding the same digit: eight ballots
        uint64_t same = __ballot(valid);
        for (int b = 0; b < 8; b++) {
            const uint64_t bal = __ballot(((d >> b) & 1) != 0);
            same &= ((d >> b) & 1) ? bal : ~bal;
        }
        const int rank = __popcll(same & below);
        if (valid && rank == 0) wcnt[wave][d] = (uint32_t)__popcll(same);
        __syncthreads();
        if (valid) {
            uint32_t off = next[d] + (uint32_t)rank;
            for (int w = 0; w < wave; w++) off += wcnt[w][d];
            if (off < n) {
                kout[off] = key;
                if (vin) vout[off] = vin[idx];
            }
        }
        __syncthreads();
        uint32_t add = 0;
        for (int w = 0; w < RX_WAVES; w++) add += wcnt[w][t];
        next[t] += add;
        __syncthreads();
    }
}

extern "C" size_t rv_radix_scratch_bytes(size_t n, int with_values) {
    const size_t tiles = (n + RX_TILE - 1) / RX_TILE;
    const size_t m = tiles * RX_BINS;
    size_t bytes = n * 8 + (with_values ? n * 4 : 0);
    bytes = (bytes + 255) / 256 * 256;
    bytes += (m + rx_scan_words(m)) * 4 + 256;
    return bytes;
}

extern "C" int rv_radix_sort(const uint64_t *keys_in, uint64_t *keys_out,
                             const uint32_t *vals_in, uint32_t *vals_out, size_t n,
                             int begin_bit, int end_bit, void *scratch, size_t scratch_bytes,
                             void *stream) {
    if (!keys_in || !keys_out || keys_in == keys_out || (vals_in && !vals_out) ||
        (vals_in && vals_in == vals_out) || begin_bit < 0 ||
        end_bit > 64 || begin_bit >= end_bit || n > 0xffffffffull || !scratch ||
        scratch_bytes < rv_radix_scratch_bytes(n, vals_in != nullptr)) {
        rv_set_error("rv_radix_sort: bad arguments");
        return -1;
    }
    hipStream_t s = (hipStream_t)stream;
    if (n == 0) return 0;
    const int passes = (end_bit - begin_bit + 7) / 8;
    const size_t tiles = (n + RX_TILE - 1) / RX_TILE;
    const size_t m = tiles * RX_BINS;
    unsigned char *p = (unsigned char *)scratch;
    uint64_t *kt = (uint64_t *)p;
    uint32_t *vt = vals_in ? (uint32_t *)(p + n * 8) : nullptr;
    size_t off = n * 8 + (vals_in ? n * 4 : 0);
    off = (off + 255) / 256 * 256;
    uint32_t *counts = (uint32_t *)(p + off);
    uint32_t *work = counts + m;
    // pass i reads src, writes dst; the last pass writes the caller's output
    const uint64_t *ksrc = keys_in;
    const uint32_t *vsrc = vals_in;
    for (int i = 0; i < passes; i++) {
        const bool to_out = ((passes - 1 - i) % 2) == 0;
        uint64_t *kdst = to_out ? keys_out : kt;
        uint32_t *vdst = vals_in ? (to_out ? vals_out : vt) : nullptr;
        const int shift = begin_bit + 8 * i;
        const int bits = end_bit - shift < 8 ? end_bit - shift : 8;
        const uint32_t mask = (1u << bits) - 1u;
        rx_count_kernel<<<(unsigned)tiles, RX_THREADS, 0, s>>>(ksrc, n, shift, mask, counts,
                                                               (int)tiles);
        if (rx_check(hipGetLastError(), "rx_count_kernel") || rx_scan(counts, m, work, s))
            return -1;
        rx_scatter_kernel<<<(unsigned)tiles, RX_THREADS, 0, s>>>(ksrc, kdst, vsrc, vdst, n, shift,
                                                                 mask, counts, (int)tiles);
        if (rx_check(hipGetLastError(), "rx_scatter_kernel")) return -1;
        ksrc = kdst;
        vsrc = vdst;
    }
    return 0;
}
