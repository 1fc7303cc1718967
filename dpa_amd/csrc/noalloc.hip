// noalloc.hip -- the allocation-free form of dpah_advect_endpoints_device.
//
// For callers that cannot allocate device memory (code under hipGraph
// capture, tight serving loops, harnesses that forbid hipMalloc).  The
// endpoint launch's scratch -- the step counter, the persistent queue's part
// heads and the queue-ordering keys, order table and histogram -- lives in one
// file-scope device array instead of the context's counters and its caching
// allocator.  Every GPU the library runs on holds its own copy of the array
// until the process ends; dpah_create looks up its address on that device once
// (hipGetSymbolAddress), and every use below goes through that address, never
// through the array's name (which is only a host-side stand-in).  The kernels
// are the ones dpah_advect_endpoints_device launches, so the results are the
// same bit for bit.
#include <mutex>

#include "order_kernels.hpp"
#include "runtime.hpp"

namespace dpah {

namespace {

// Layout of the scratch array (byte offsets): counters (kCounterCount u64:
// [1] particle-steps, [kQueueHeads..] queue heads), then the ordering's key
// and order tables (one u32 per particle each), then its histogram and tile
// sums (at most 2^(3 kOrderMaxBits) + 1 keys).
constexpr uint64_t kCountersBytes = 256;
constexpr uint64_t kMaxParticles = DPAH_NOALLOC_MAX_PARTICLES;
constexpr uint64_t kMaxKeys = (1ull << (3 * kOrderMaxBits)) + 1;
constexpr uint64_t kMaxTiles = (kMaxKeys + kOrderTile - 1) / kOrderTile;
constexpr uint64_t kCountWords = kMaxKeys + kMaxTiles;
constexpr uint64_t kKeyOffset = kCountersBytes;
constexpr uint64_t kOrderOffset = kKeyOffset + kMaxParticles * sizeof(uint32_t);
constexpr uint64_t kCountOffset = kOrderOffset + kMaxParticles * sizeof(uint32_t);
constexpr uint64_t kScratchBytes = (kCountOffset + kCountWords * sizeof(uint32_t) + 255) / 256 * 256;
static_assert(kCountersBytes >= kCounterCount * sizeof(unsigned long long), "counters fit");
static_assert(kScratchBytes <= DPAH_NOALLOC_SCRATCH_BYTES, "scratch size stated in the header");

static __device__ __align__(256) unsigned char scratch[DPAH_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_ptr[kMaxDevices] = {};   // the array's address on each device

}  // namespace

// Called by dpah_create with `device` current: looks up this device's copy of
// the scratch array once per process.
void noalloc_setup(int device) {
  DPAH_REQUIRE(device >= 0 && device < kMaxDevices, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return;
  void* p = nullptr;
  DPAH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  DPAH_REQUIRE(p != nullptr, "scratch array has no device address");
  g_ptr[device] = static_cast<unsigned char*>(p);
}

}  // namespace dpah

using namespace dpah;

DPAH_EXPORT int dpah_advect_endpoints_device_noalloc(
    dpah_context* ctx, const dpah_field* field, const float* seeds, const uint64_t* remaining,
    uint64_t uniform_remaining, uint64_t count, float step_size, float start_time, int integrator,
    float* positions_out, uint64_t* remaining_out, uint8_t* status_out, uint64_t* total_steps) {
  return guarded([&] {
    DPAH_REQUIRE(ctx && field, "null argument");
    DPAH_REQUIRE(field->ctx == ctx, "field belongs to another context");
    DPAH_REQUIRE(count == 0 || (seeds && positions_out), "seeds / positions_out is null");
    DPAH_REQUIRE(remaining || uniform_remaining <= (1ull << 32) - 3,
                 "remaining iterations must be below 2^32 - 2");
    DPAH_REQUIRE(integrator == DPAH_INTEGRATOR_DPA_RK4 || integrator == DPAH_INTEGRATOR_RK4,
                 "unknown integrator");
    DPAH_REQUIRE(count <= kMaxParticles,
                 "count exceeds DPAH_NOALLOC_MAX_PARTICLES (2^21), the particles the fixed "
                 "scratch array holds; use dpah_advect_endpoints_device");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    unsigned char* base = nullptr;
    {
      std::lock_guard<std::mutex> g(g_mu);
      base = g_ptr[ctx->device];
    }
    DPAH_REQUIRE(base, "scratch array not set up on this device");
    unsigned long long* counters = reinterpret_cast<unsigned long long*>(base);
    OrderScratch os;
    os.key = reinterpret_cast<uint32_t*>(base + kKeyOffset);
    os.order = reinterpret_cast<uint32_t*>(base + kOrderOffset);
    os.count = reinterpret_cast<uint32_t*>(base + kCountOffset);
    os.n_cap = kMaxParticles;
    os.count_cap = kCountWords;
    AdvectLaunch L{};
    L.f = &field->view;
    L.seeds = seeds;
    L.rem_in = remaining;
    L.uniform_rem = uniform_remaining;
    L.pos_out = positions_out;
    L.rem_out = remaining_out;
    L.status_out = status_out;
    L.n = count;
    L.dt = step_size;
    L.t0 = start_time;
    L.integrator = integrator;
    L.persistent = true;
    L.order_scratch = &os;
    L.counters = counters;
    // kernel timing records events, which the runtime may have to create:
    // this form launches untimed
    const bool timing = ctx->timing;
    ctx->timing = false;
    try {
      launch_advect(ctx, L);
    } catch (...) {
      ctx->timing = timing;
      throw;
    }
    ctx->timing = timing;
    if (total_steps) {
      unsigned long long s = 0;
      DPAH_HIP(hipMemcpyAsync(&s, counters + 1, sizeof(s), hipMemcpyDeviceToHost, ctx->stream));
      DPAH_HIP(hipStreamSynchronize(ctx->stream));
      *total_steps = s;
    }
  });
}
