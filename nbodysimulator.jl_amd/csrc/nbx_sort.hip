// nbx_sort.hip -- spatial order of the mixed path's state, computed on the device.
//
// A Lennard-Jones list run keeps its atoms sorted by cell of the list grid (z-major; the
// atoms of a cell by user index) and re-sorts them every few thousand steps, so that the
// neighbours gathered by the list force kernel stay near each other in memory as the
// liquid mixes.  The new order comes from one radix sort of 64-bit keys
// cell * n + user index (unique, so the order is a total one and reproducible), with the
// current slot as the value; no data crosses to the host.
#include "nbx_mixed.h"

#include <rocprim/device/device_radix_sort.hpp>

// key of slot s: the cell of its position times n, plus the user index of its atom
__global__ void __launch_bounds__(256)
k_sort_keys(const double* __restrict__ x, int64_t n, double L, int g,
            const int* __restrict__ perm, unsigned long long* __restrict__ keys,
            int* __restrict__ vals) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    int64_t cell = 0;
    for (int k = 2; k >= 0; --k) {
        const double xk = x[k * n + s];
        const double w = xk - L * floor(xk / L);
        int ci = (int)(w / L * g);
        ci = ci < 0 ? 0 : (ci >= g ? g - 1 : ci);
        cell = cell * g + ci;
    }
    keys[s] = (unsigned long long)(cell * n + perm[s]);
    vals[s] = (int)s;
}

// the new order: slot t holds old slot map[t], i.e. the user's atom perm[map[t]]
__global__ void __launch_bounds__(256)
k_sort_maps(const int* __restrict__ map, const int* __restrict__ perm, int64_t n,
            int* __restrict__ perm_new, int* __restrict__ iperm_new) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int u = perm[map[t]];
    perm_new[t] = u;
    iperm_new[u] = (int)t;
}

__global__ void __launch_bounds__(256) k_iota(int* __restrict__ p, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = (int)t;
}

static int key_bits(unsigned long long top) {
    int b = 1;
    while (b < 64 && (top >> b) != 0ull) ++b;
    return b;
}

static int sort_bits(int64_t n, int g) {
    return key_bits((unsigned long long)g * g * g * (unsigned long long)n);
}

size_t mx_sort_temp_bytes(int64_t n, int g) {
    size_t bytes = 0;
    if (rocprim::radix_sort_pairs(nullptr, bytes, (unsigned long long*)nullptr,
                                  (unsigned long long*)nullptr, (int*)nullptr, (int*)nullptr,
                                  (size_t)n, 0, sort_bits(n, g), (hipStream_t)0) != hipSuccess)
        return 0;
    return bytes;
}

// descents of the key sequence (slots s with key[s] > key[s + 1]): 0 for a state in
// cell order, about n / 80 for an FCC lattice in generation order, about n / 2 shuffled
__global__ void __launch_bounds__(256)
k_sort_descents(const unsigned long long* __restrict__ keys, int64_t n,
                unsigned* __restrict__ count) {
    __shared__ unsigned part;
    if (threadIdx.x == 0) part = 0;
    __syncthreads();
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s + 1 < n && keys[s] > keys[s + 1]) atomicAdd(&part, 1u);
    __syncthreads();
    if (threadIdx.x == 0 && part != 0) atomicAdd(count, part);
}

__global__ void k_sort_zero(unsigned* p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) p[0] = 0;
}

int mx_launch_sort_keys(hipStream_t s, const double* x, int64_t n, double L, int g,
                        const int* perm, unsigned long long* keys, int* vals,
                        unsigned* descents) {
    if (n <= 0) return 0;
    const unsigned blocks = nbx_blocks(n, 256);
    hipLaunchKernelGGL(k_sort_zero, dim3(1), dim3(64), 0, s, descents);
    hipLaunchKernelGGL(k_sort_keys, dim3(blocks), dim3(256), 0, s, x, n, L, g, perm, keys,
                       vals);
    hipLaunchKernelGGL(k_sort_descents, dim3(blocks), dim3(256), 0, s, keys, n, descents);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mx_launch_sort_apply(hipStream_t s, int64_t n, int g, unsigned long long* keys_in,
                         unsigned long long* keys_out, int* vals_in, int* map, void* temp,
                         size_t temp_bytes, const int* perm, int* perm_new, int* iperm_new) {
    if (n <= 0) return 0;
    const unsigned blocks = nbx_blocks(n, 256);
    size_t bytes = temp_bytes;
    if (rocprim::radix_sort_pairs(temp, bytes, keys_in, keys_out, vals_in, map, (size_t)n, 0,
                                  sort_bits(n, g), s) != hipSuccess)
        return -1;
    hipLaunchKernelGGL(k_sort_maps, dim3(blocks), dim3(256), 0, s, map, perm, n, perm_new,
                       iperm_new);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

void mx_launch_iota(hipStream_t s, int* p, int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_iota, dim3(nbx_blocks(n, 256)), dim3(256), 0, s, p, n);
}

// the per-atom state reordered: new[s] = old[map[s]] for every array, through a scratch
// buffer of (3 n3 + n1) n doubles (two launches for the whole state)
__global__ void __launch_bounds__(256)
k_perm_gather(const int* __restrict__ map, int64_t n, mx_perm_arrays A, double* __restrict__ tmp) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int64_t j = map[s];
    int o = 0;
    for (int q = 0; q < A.n3; ++q, o += 3)
#pragma unroll
        for (int k = 0; k < 3; ++k) tmp[(o + k) * n + s] = A.a3[q][k * n + j];
    for (int q = 0; q < A.n1; ++q, ++o) tmp[o * n + s] = A.a1[q][j];
}

__global__ void __launch_bounds__(256)
k_perm_store(int64_t n, mx_perm_arrays A, const double* __restrict__ tmp) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    int o = 0;
    for (int q = 0; q < A.n3; ++q, o += 3)
#pragma unroll
        for (int k = 0; k < 3; ++k) A.a3[q][k * n + s] = tmp[(o + k) * n + s];
    for (int q = 0; q < A.n1; ++q, ++o) A.a1[q][s] = tmp[o * n + s];
}

void mx_launch_permute(hipStream_t s, const int* map, int64_t n, const mx_perm_arrays& A,
                       double* tmp) {
    if (n <= 0) return;
    const unsigned blocks = nbx_blocks(n, 256);
    hipLaunchKernelGGL(k_perm_gather, dim3(blocks), dim3(256), 0, s, map, n, A, tmp);
    hipLaunchKernelGGL(k_perm_store, dim3(blocks), dim3(256), 0, s, n, A, tmp);
}
