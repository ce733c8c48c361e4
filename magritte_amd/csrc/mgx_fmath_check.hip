/*
 * Device evaluation of the restated float functions (mgx_fmath_ref.h) over
 * whole ranges of float arguments, reduced to per-chunk digests, so that the
 * GPU build can be compared with the C library on every argument without
 * moving 2^32 results to the host (mgx_fmath_digest in mgx.h).
 */
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mgx.h"
#include "mgx_error.h"
#include "mgx_fmath_ref.h"

namespace {

constexpr uint32_t kBlock = 256;

__global__ void zero_kernel(uint64_t *d, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = 0;
}

/* grid: (blocks per chunk, chunks); each thread sums the digests of its
   arguments, the block reduces in LDS and adds into its chunk's digest */
__global__ __launch_bounds__(kBlock) void digest_kernel(int fn, float y, uint64_t first,
                                                         uint32_t chunk_log2, uint64_t *digest)
{
    mgx_ref_tables_init();
    __shared__ uint64_t s[kBlock];
    const uint64_t chunk = blockIdx.y;
    const uint64_t base = first + (chunk << chunk_log2);
    const uint64_t n = 1ull << chunk_log2;
    uint64_t acc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * kBlock) {
        const uint32_t xb = (uint32_t)(base + i);
        acc += mgx_fm_digest(xb, mgx_ref_eval_bits(fn, xb, y));
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t h = kBlock / 2; h > 0; h /= 2) {
        if (threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(reinterpret_cast<unsigned long long *>(&digest[chunk]),
                                    (unsigned long long)s[0]);
}

} // namespace

extern "C" int mgx_fmath_digest(int fn, float y, uint64_t first, uint32_t nchunks,
                                uint32_t chunk_log2, uint64_t *digest, void *stream)
{
    mgx::clear_error();
    if (fn < MGX_FN_EXPF || fn > MGX_FN_POWF) return mgx::fail(MGX_E_ARG, "unknown function %d", fn);
    if (!digest || nchunks == 0) return mgx::fail(MGX_E_ARG, "null digest or no chunks");
    if (chunk_log2 < 8 || chunk_log2 > 32) return mgx::fail(MGX_E_ARG, "chunk_log2 out of [8, 32]");
    if (first + ((uint64_t)nchunks << chunk_log2) > (1ull << 32))
        return mgx::fail(MGX_E_ARG, "the range goes beyond the 2^32 float arguments");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(zero_kernel, dim3((nchunks + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                       digest, nchunks);
    const uint64_t per = (1ull << chunk_log2) / kBlock;
    const uint32_t bx = (uint32_t)(per < 1024 ? (per ? per : 1) : 1024);
    for (uint32_t c0 = 0; c0 < nchunks; c0 += 65535) {
        const uint32_t nc = nchunks - c0 < 65535 ? nchunks - c0 : 65535;
        hipLaunchKernelGGL(digest_kernel, dim3(bx, nc), dim3(kBlock), 0, st, fn, y,
                           first + ((uint64_t)c0 << chunk_log2), chunk_log2, digest + c0);
    }
    const hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) return mgx::fail(MGX_E_HIP, "%s", hipGetErrorString(e));
    return MGX_OK;
}
