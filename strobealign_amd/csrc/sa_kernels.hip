// HIP kernels of the mapping path: seeding -> index lookups/hits -> anchors ->
// collinear chaining -> NAMs, and the extension kernel. Each thread runs one
// work item of sa_items.hpp (a read, a read orientation or an extension request).
#include <hip/hip_runtime.h>

#include "sa_items.hpp"

namespace sa {

__global__ void __launch_bounds__(256) k_seed(const uint8_t* __restrict__ seq, const uint64_t* __restrict__ off,
                                             uint32_t n, SeedParams p, Syncmer* __restrict__ sy,
                                             QRand* __restrict__ qr, uint32_t* __restrict__ qr_n) {
    item_seed(blockIdx.x * blockDim.x + threadIdx.x, seq, off, n, p, sy, qr, qr_n);
}

__global__ void __launch_bounds__(256) k_hits(IndexView ix, MapParams mp, const uint64_t* __restrict__ off,
                                             uint32_t n, const QRand* __restrict__ qr,
                                             const uint32_t* __restrict__ qr_n, Hit* __restrict__ hits,
                                             uint64_t* __restrict__ cnt, HitCounts* __restrict__ hc,
                                             uint64_t* __restrict__ acap) {
    item_hits(blockIdx.x * blockDim.x + threadIdx.x, ix, mp, off, n, qr, qr_n, hits, cnt, hc, acap);
}

__global__ void __launch_bounds__(256) k_chain(IndexView ix, ChainParams cp, const uint64_t* __restrict__ off,
                                              uint32_t n, const Hit* __restrict__ hits,
                                              const HitCounts* __restrict__ hc, const uint64_t* __restrict__ acap,
                                              const uint64_t* __restrict__ aoff, ChainBufs b,
                                              uint32_t* __restrict__ nam_n, uint32_t* __restrict__ anchors_raw,
                                              const float* __restrict__ log2tab,
                                              const float* __restrict__ same_diag) {
    item_chain(blockIdx.x * blockDim.x + threadIdx.x, ix, cp, off, n, hits, hc, acap, aoff, b, nam_n, anchors_raw,
               log2tab, same_diag);
}

__global__ void __launch_bounds__(256) k_finalize(const uint8_t* __restrict__ seq, const uint64_t* __restrict__ off,
                                                 uint32_t n, RefView ref, uint32_t k,
                                                 const uint32_t* __restrict__ nam_n,
                                                 const uint64_t* __restrict__ aoff, const NamRec* __restrict__ nams_in,
                                                 const uint64_t* __restrict__ out_off, NamRec* __restrict__ nams_out) {
    item_finalize(blockIdx.x * blockDim.x + threadIdx.x, seq, off, n, ref, k, nam_n, aoff, nams_in, out_off, nams_out);
}

__global__ void __launch_bounds__(128) k_extend(const uint8_t* __restrict__ seq, const uint64_t* __restrict__ off,
                                               RefView ref, Scores sc, uint32_t k,
                                               const ExtRequest* __restrict__ req, uint32_t nreq,
                                               ExtResult* __restrict__ res, uint32_t* __restrict__ cigars,
                                               const uint32_t* __restrict__ chain_anchors, uint8_t* __restrict__ scratch,
                                               uint64_t scratch_per, uint32_t first) {
    item_extend(blockIdx.x * blockDim.x + threadIdx.x, seq, off, ref, sc, k, req, nreq, res, cigars, chain_anchors,
                scratch, scratch_per, first);
}

}  // namespace sa
