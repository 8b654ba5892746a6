// One iteration's vector work of the conjugate-gradient loop of the
// normal-equation solves (FP64), after q = A^T A p has been applied:
//   pq = p.q, alpha = rr / pq (0 when pq = 0), x += alpha p, r -= alpha q,
//   rr' = r.r, p = r + (rr' / rr) p.
// The CG loop of the reference's own solve is torch vector algebra
// (main_CAGR.py:94-144: alpha, beta and the updates as tensor expressions);
// here the same recurrence runs in three launches instead of a dozen torch
// kernels, each reduction in a fixed order (fixed block ranges, fixed trees
// inside a block, the block partials added in block order by every block that
// needs the total), so the iterates do not depend on scheduling.
#include "common.h"

namespace cagr {
namespace {

constexpr int kCgThreads = 256;
constexpr int kCgBlocks = 512;   // block partials per reduction

// sum over the block of one value per thread, in a fixed tree (lane
// butterflies, then the waves' sums in order); every thread gets the sum
__device__ inline double block_sum(double v, double* lds) {
    for (int s = 32; s > 0; s >>= 1) v += __shfl_xor(v, s, 64);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds[wave] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < kCgThreads / 64; ++w) t += lds[w];
    __syncthreads();
    return t;
}

// the sum of the kCgBlocks partials, in block order, by every block
__device__ inline double total_of(const double* __restrict__ part, double* lds) {
    double v = 0.0;
    for (int b = threadIdx.x; b < kCgBlocks; b += kCgThreads) v += part[b];
    return block_sum(v, lds);
}

// block b handles elements [b * chunk, (b + 1) * chunk) in a strided loop
__device__ inline void range_of(int64_t len, int64_t& lo, int64_t& hi) {
    const int64_t chunk = (len + kCgBlocks - 1) / kCgBlocks;
    lo = (int64_t)blockIdx.x * chunk;
    hi = lo + chunk < len ? lo + chunk : len;
}

__global__ __launch_bounds__(kCgThreads) void k_cg_pq(int64_t len, const double* __restrict__ p,
                                                      const double* __restrict__ q,
                                                      double* __restrict__ part) {
    __shared__ double lds[kCgThreads / 64];
    int64_t lo, hi;
    range_of(len, lo, hi);
    double v = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += kCgThreads) v = fma(p[i], q[i], v);
    v = block_sum(v, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ __launch_bounds__(kCgThreads) void k_cg_xr(int64_t len, double* __restrict__ x,
                                                      double* __restrict__ r,
                                                      const double* __restrict__ p,
                                                      const double* __restrict__ q,
                                                      const double* __restrict__ rr,
                                                      const double* __restrict__ part_pq,
                                                      double* __restrict__ part_rr) {
    __shared__ double lds[kCgThreads / 64];
    const double pq = total_of(part_pq, lds);
    const double alpha = pq != 0.0 ? rr[0] / pq : 0.0;
    int64_t lo, hi;
    range_of(len, lo, hi);
    double v = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += kCgThreads) {
        x[i] = fma(alpha, p[i], x[i]);
        const double ri = fma(-alpha, q[i], r[i]);
        r[i] = ri;
        v = fma(ri, ri, v);
    }
    v = block_sum(v, lds);
    if (threadIdx.x == 0) part_rr[blockIdx.x] = v;
}

__global__ __launch_bounds__(kCgThreads) void k_cg_p(int64_t len, const double* __restrict__ r,
                                                     double* __restrict__ p,
                                                     const double* __restrict__ rr,
                                                     const double* __restrict__ part_rr,
                                                     double* __restrict__ rr_new) {
    __shared__ double lds[kCgThreads / 64];
    const double rn = total_of(part_rr, lds);
    const double beta = rr[0] != 0.0 ? rn / rr[0] : 0.0;
    int64_t lo, hi;
    range_of(len, lo, hi);
    for (int64_t i = lo + threadIdx.x; i < hi; i += kCgThreads) p[i] = fma(beta, p[i], r[i]);
    if (blockIdx.x == 0 && threadIdx.x == 0) rr_new[0] = rn;
}

}  // namespace
}  // namespace cagr

using namespace cagr;

extern "C" int cagr_cg_step(int64_t len, double* x, double* r, double* p, const double* q,
                            const double* rr, double* rr_new, double* work, int64_t work_len,
                            void* stream) {
    CAGR_REQUIRE(len >= 0, CAGR_ERR_ARG, "cg step: negative length");
    CAGR_REQUIRE(work_len >= 2 * kCgBlocks, CAGR_ERR_ARG,
                 "cg step: the workspace needs cagr_cg_work_len() doubles");
    CAGR_REQUIRE(rr != rr_new, CAGR_ERR_ARG, "cg step: rr and rr_new must differ");
    CAGR_REQUIRE(x && r && p && q && rr && rr_new && work, CAGR_ERR_ARG, "cg step: null pointer");
    const hipStream_t st = (hipStream_t)stream;
    double* part_pq = work;
    double* part_rr = work + kCgBlocks;
    k_cg_pq<<<kCgBlocks, kCgThreads, 0, st>>>(len, p, q, part_pq);
    CAGR_LAUNCHED();
    k_cg_xr<<<kCgBlocks, kCgThreads, 0, st>>>(len, x, r, p, q, rr, part_pq, part_rr);
    CAGR_LAUNCHED();
    k_cg_p<<<kCgBlocks, kCgThreads, 0, st>>>(len, r, p, rr, part_rr, rr_new);
    CAGR_LAUNCHED();
    return CAGR_OK;
}

extern "C" int64_t cagr_cg_work_len(void) { return 2 * kCgBlocks; }
