// Allocation-free batched run (octmps_run_bscan_noalloc): octmps_run_bscan with every device
// buffer of the run (region and set tables, scan slots, photon-id segments, tallies, control
// words, per-lane continuation photons) carved from one file-scope device array instead of
// the context's growable work buffers, for callers that must not allocate device memory.
//
// The array is `scratch`, kScratchBytes per GPU, held until the process ends. Its device
// address is looked up once per context (octmps_ctx_create, with the context's device
// current) and kept in ctx->scratch; host code passes that address to kernels and copies,
// never the array's name. Every call on a GPU shares the array and leaves its last contents
// behind: the run clears the tallies, control words and completion counters with fill
// kernels before the transport launch, and calls that overlap on one GPU (two contexts on
// the same device) race on it.
#include <hip/hip_runtime.h>

#include "device_types.hpp"
#include "host_common.hpp"
#include "runtime_internal.hpp"

namespace octmps {

constexpr size_t kScratchBytes = 64ull << 20;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

octmps_status scratch_init(octmps_ctx *ctx) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "scratch address");
    ctx->scratch = ptr;
    ctx->scratch_bytes = kScratchBytes;
    return OCTMPS_OK;
}

}  // namespace octmps

extern "C" octmps_status octmps_run_bscan_noalloc(octmps_ctx *c, int32_t mesh_id, const octmps_params *p,
                                                  int32_t n, const octmps_ascan *scans, uint64_t seed,
                                                  const octmps_batch_opts *opts, uint64_t *tallies_out,
                                                  octmps_run_stats *stats) {
    if (!c) return octmps::fail(OCTMPS_E_INVALID, "null argument");
    if (!c->scratch) return octmps::fail(OCTMPS_E_DEVICE, "device scratch not set up");
    return octmps::run_scans(c, mesh_id, p, n, scans, seed, opts, tallies_out, stats, nullptr, 0, c->scratch,
                             c->scratch_bytes);
}
