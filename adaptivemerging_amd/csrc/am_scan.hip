// am_scan.hip - device-wide exclusive scan (64-wide wavefront shuffles, LDS
// for wave totals) used for every count -> offset compaction, and the key sort
// of broad-phase pairs and colour-ordered groups.
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>

#include "am_internal.hpp"

namespace am {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;  // 1024 elements per block

// inclusive scan across the 64 lanes of a wavefront
AMD_INLINE int wave_inclusive(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// exclusive scan of the block's thread values; returns the block total
AMD_INLINE int block_exclusive(int v, int* total) {
    __shared__ int wave_sums[kScanThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int inc = wave_inclusive(v);
    if (lane == 63) wave_sums[wave] = inc;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int q = 0; q < kScanThreads / 64; q++) {
        int s = wave_sums[q];
        if (q < wave) off += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_tile(const int* in, int* out, int n, int* block_sums,
                                                            int* d_total) {
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
    int sum = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        int idx = base + i;
        v[i] = idx < n ? in[idx] : 0;
        sum += v[i];
    }
    int total;
    int run = block_exclusive(sum, &total);
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        int idx = base + i;
        if (idx < n) out[idx] = run;
        run += v[i];
    }
    if (threadIdx.x == 0) {
        if (block_sums) block_sums[blockIdx.x] = total;
        if (d_total) *d_total = total;
    }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(int* out, int n, const int* block_offsets) {
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    const int add = block_offsets[blockIdx.x];
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        int idx = base + i;
        if (idx < n) out[idx] += add;
    }
}

__global__ void k_zero_one(int* p) { *p = 0; }

static int n_tiles(int n) { return (n + kScanTile - 1) / kScanTile; }

size_t scan_temp_bytes(int n) {
    size_t bytes = 0;
    while (n > kScanTile) {
        int t = n_tiles(n);
        bytes += 2 * (size_t)t * sizeof(int) + 256;
        n = t;
    }
    return bytes + 256;
}

// Recursive: tile scan, scan of tile totals, add.
void exclusive_scan(hipStream_t s, const int* in, int* out, int n, int* d_total, void* temp) {
    if (n <= 0) {
        if (d_total) hipLaunchKernelGGL(k_zero_one, dim3(1), dim3(1), 0, s, d_total);
        return;
    }
    if (n <= kScanTile) {
        hipLaunchKernelGGL(k_scan_tile, dim3(1), dim3(kScanThreads), 0, s, in, out, n, (int*)nullptr, d_total);
        return;
    }
    int t = n_tiles(n);
    int* sums = (int*)temp;
    int* sums_scanned = sums + t;
    char* next = (char*)(sums_scanned + t);
    next += (256 - ((uintptr_t)next & 255)) & 255;
    hipLaunchKernelGGL(k_scan_tile, dim3(t), dim3(kScanThreads), 0, s, in, out, n, sums, (int*)nullptr);
    exclusive_scan(s, sums, sums_scanned, t, d_total, next);
    hipLaunchKernelGGL(k_scan_add, dim3(t), dim3(kScanThreads), 0, s, out, n, sums_scanned);
}

size_t sort_temp_bytes(int n) {
    size_t bytes = 0;
    rocprim::radix_sort_keys((void*)nullptr, bytes, (uint64_t*)nullptr, (uint64_t*)nullptr, (size_t)n, 0, 64);
    return bytes;
}

void sort_u64(hipStream_t s, uint64_t* keys_in, uint64_t* keys_out, int n, void* temp, size_t temp_bytes,
              int begin_bit, int end_bit) {
    if (n <= 0) return;
    size_t bytes = temp_bytes;
    rocprim::radix_sort_keys(temp, bytes, keys_in, keys_out, (size_t)n, (unsigned)begin_bit, (unsigned)end_bit, s);
}

}  // namespace am
