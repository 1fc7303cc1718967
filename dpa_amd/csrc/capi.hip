// capi.hip -- the C ABI declared in include/dpa_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "runtime.hpp"


namespace dpah {

std::string& error_slot() {
  thread_local std::string msg;
  return msg;
}

void set_error(const std::string& msg) { error_slot() = msg; }

}  // namespace dpah

using namespace dpah;

// ------------------------------------------------------------ context ----

namespace {
uint64_t round_block(uint64_t bytes) {
  if (bytes == 0) bytes = 1;
  const uint64_t g = bytes >= (1ull << 20) ? (2ull << 20) : 512;
  return (bytes + g - 1) / g * g;
}
}  // namespace

void dpah_context::trim_cache() {
  for (auto& kv : cached) {
    (void)hipFree(kv.second);
    block_size.erase(kv.second);
  }
  cached.clear();
  cached_bytes = 0;
}

void dpah_context::drain_deferred() {
  std::vector<dpah_result*> take;
  {
    std::lock_guard<std::mutex> lk(deferred_mu);
    take.swap(deferred);
  }
  for (dpah_result* r : take) delete r;
}

void* dpah_context::host_staging(uint64_t bytes) {
  if (bytes > kPinnedStagingMax) return nullptr;
  if (bytes > pinned_bytes) {
    if (pinned) {
      DPAH_HIP(hipStreamSynchronize(stream));
      (void)hipHostFree(pinned);
      pinned = nullptr;
      pinned_bytes = 0;
    }
    // 2 MiB granules, at least 8 MiB, so that small reads never regrow it
    const uint64_t want = std::max<uint64_t>((bytes + (2ull << 20) - 1) & ~((2ull << 20) - 1),
                                             8ull << 20);
    DPAH_HIP(hipHostMalloc(&pinned, want, hipHostMallocDefault));
    pinned_bytes = want;
  }
  return pinned;
}

void dpah_context::download(void* dst, const void* src, uint64_t bytes, hipStream_t s) {
  if (!bytes) return;
  hipPointerAttribute_t at{};
  bool pinned_dst = false;
  if (hipPointerGetAttributes(&at, dst) == hipSuccess) pinned_dst = at.type == hipMemoryTypeHost;
  else (void)hipGetLastError();   // pageable memory is not known to the runtime
  uint8_t* stage = nullptr;
  if (!pinned_dst && bytes >= kStagedMin)
    stage = static_cast<uint8_t*>(host_staging(2 * kStagedChunk));
  if (!stage) {
    DPAH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    DPAH_HIP(hipStreamSynchronize(s));
    return;
  }
  hipEvent_t ev[2] = {take_event(), take_event()};
  const uint64_t n = (bytes + kStagedChunk - 1) / kStagedChunk;
  auto issue = [&](uint64_t i) {
    const uint64_t off = i * kStagedChunk, len = std::min(kStagedChunk, bytes - off);
    DPAH_HIP(hipMemcpyAsync(stage + (i & 1) * kStagedChunk, static_cast<const uint8_t*>(src) + off,
                            len, hipMemcpyDeviceToHost, s));
    DPAH_HIP(hipEventRecord(ev[i & 1], s));
  };
  try {
    issue(0);
    for (uint64_t i = 0; i < n; ++i) {
      // chunk i + 1 goes into the buffer chunk i - 1 left, already copied out
      if (i + 1 < n) issue(i + 1);
      DPAH_HIP(hipEventSynchronize(ev[i & 1]));
      const uint64_t off = i * kStagedChunk, len = std::min(kStagedChunk, bytes - off);
      uint8_t* to = static_cast<uint8_t*>(dst) + off;
      const uint8_t* from = stage + (i & 1) * kStagedChunk;
      const uint64_t part = ((len + kStagedThreads - 1) / kStagedThreads + 63) & ~uint64_t(63);
      std::thread w[kStagedThreads - 1];
      for (int k = 1; k < kStagedThreads; ++k) {
        const uint64_t a = std::min(len, k * part), b = std::min(len, a + part);
        if (a >= b) continue;
        try {
          w[k - 1] = std::thread([=] { std::memcpy(to + a, from + a, b - a); });
        } catch (...) {   // no thread to be had: this part on the calling thread
          std::memcpy(to + a, from + a, b - a);
        }
      }
      std::memcpy(to, from, std::min(len, part));
      for (auto& t : w)
        if (t.joinable()) t.join();
    }
  } catch (...) {
    (void)hipStreamSynchronize(s);
    event_pool.push_back(ev[0]);
    event_pool.push_back(ev[1]);
    throw;
  }
  event_pool.push_back(ev[0]);
  event_pool.push_back(ev[1]);
}

void* dpah_context::alloc(uint64_t bytes) {
  drain_deferred();
  const uint64_t want = round_block(bytes);
  if (caching) {
    auto it = cached.lower_bound(want);
    if (it != cached.end() && it->first <= 2 * want) {
      void* p = it->second;
      bytes_in_use += it->first;
      cached_bytes -= it->first;
      cached.erase(it);
      return p;
    }
  }
  if (bytes_in_use.load() + want > memory_budget)
    throw Error(DPAH_ERR_NOMEM, "device memory budget exceeded: " + std::to_string(want) +
                                    " bytes requested, " + std::to_string(bytes_in_use.load()) +
                                    " of " + std::to_string(memory_budget) + " in use");
  if (bytes_in_use.load() + cached_bytes + want > memory_budget) trim_cache();
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess && cached_bytes) {
    (void)hipGetLastError();
    trim_cache();
    e = hipMalloc(&p, want);
  }
  if (e != hipSuccess)
    throw Error(DPAH_ERR_NOMEM, std::string("hipMalloc(") + std::to_string(want) + "): " +
                                    hipGetErrorString(e));
  block_size[p] = want;
  if (!caching) direct.insert(p);
  bytes_in_use += want;
  return p;
}

void dpah_context::release(void* p, uint64_t /*bytes*/) {
  if (!p) return;
  auto it = block_size.find(p);
  if (it == block_size.end()) return;
  const uint64_t sz = it->second;
  bytes_in_use -= sz;
  auto d = direct.find(p);
  if (caching && d == direct.end()) {
    cached.emplace(sz, p);
    cached_bytes += sz;
  } else {
    // blocks taken while caching was off go back to the runtime whenever they
    // are released, so the per-request form never feeds the cache
    (void)hipFree(p);
    block_size.erase(it);
    if (d != direct.end()) direct.erase(d);
  }
}

hipEvent_t dpah_context::take_event() {
  if (!event_pool.empty()) {
    hipEvent_t e = event_pool.back();
    event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  DPAH_HIP(hipEventCreate(&e));
  return e;
}

void dpah_context::resolve_timings() {
  std::vector<dpah::PendingTiming> still;
  for (auto& t : pending) {
    if (hipEventQuery(t.stop) == hipErrorNotReady) {   // another stream, not finished yet
      still.push_back(t);
      continue;
    }
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, t.start, t.stop) == hipSuccess) {
      auto& k = timings[t.name];
      k.launches += 1;
      k.total_ms += ms;
    }
    event_pool.push_back(t.start);
    event_pool.push_back(t.stop);
  }
  pending.swap(still);
}

void dpah::materialize_host(dpah_result* r, hipStream_t s) {
  dpah_context* ctx = r->ctx;
  for (auto& round : r->rounds)
    for (auto& c : round.chunks) {
      if (!c.vertices) continue;
      // one copy of the block: vertices | colours | indices, c.bytes in all
      c.colors_off = 12 * c.V;
      c.indices_off = 16 * c.V;
      c.host.reset(new uint8_t[c.bytes], std::default_delete<uint8_t[]>());
      DPAH_HIP(hipMemcpyAsync(c.host.get(), c.vertices, c.bytes, hipMemcpyDeviceToHost, s));
    }
  if (r->d_ftle) {
    r->host_ftle.reset(new uint8_t[std::max<uint64_t>(r->ftle_count, 1) * sizeof(float)],
                       std::default_delete<uint8_t[]>());
    if (r->ftle_count)
      DPAH_HIP(hipMemcpyAsync(r->host_ftle.get(), r->d_ftle, r->ftle_count * sizeof(float),
                              hipMemcpyDeviceToHost, s));
  }
  DPAH_HIP(hipStreamSynchronize(s));
  for (auto& round : r->rounds)
    for (auto& c : round.chunks)
      if (c.vertices) {
        ctx->release(c.vertices, c.bytes);
        c.vertices = nullptr;
        c.colors = nullptr;
        c.indices = nullptr;
      }
  if (r->d_ftle) {
    ctx->release(r->d_ftle, r->ftle_count * sizeof(float));
    r->d_ftle = nullptr;
  }
}

dpah_result::~dpah_result() {
  if (d_ftle) ctx->release(d_ftle, ftle_count * sizeof(float));
  for (auto& r : rounds)
    for (auto& c : r.chunks)
      if (c.vertices) ctx->release(c.vertices, c.bytes);
}

DPAH_EXPORT const char* dpah_last_error(void) { return dpah::error_slot().c_str(); }

DPAH_EXPORT int dpah_version(void) { return 1; }

DPAH_EXPORT int dpah_create(int device, const dpah_options* options, dpah_context** out) {
  return guarded([&] {
    DPAH_REQUIRE(out, "out is null");
    int count = 0;
    DPAH_HIP(hipGetDeviceCount(&count));
    DPAH_REQUIRE(device >= 0 && device < count, "device index out of range");
    DPAH_HIP(hipSetDevice(device));
    dpah::noalloc_setup(device);
    auto* ctx = new dpah_context();
    try {
      ctx->device = device;
      hipDeviceProp_t prop;
      DPAH_HIP(hipGetDeviceProperties(&prop, device));
      ctx->cu_count = prop.multiProcessorCount;
      int xcc = 1;
      if (hipDeviceGetAttribute(&xcc, hipDeviceAttributeNumberOfXccs, device) != hipSuccess ||
          xcc < 1)
        xcc = 1;
      ctx->xcds = xcc;
      ctx->tune.parts = std::min(xcc, dpah::kMaxParts);
      size_t free_b = 0, total_b = 0;
      DPAH_HIP(hipMemGetInfo(&free_b, &total_b));
      ctx->memory_budget = (options && options->memory_budget_bytes)
                               ? options->memory_budget_bytes
                               : (uint64_t)(0.9 * (double)free_b);
      ctx->slot_budget = (options && options->slot_budget_bytes) ? options->slot_budget_bytes
                                                                 : (4ull << 30);
      ctx->timing = options && options->kernel_timing;
      ctx->arch = prop.gcnArchName;
      if (options && options->autotune_cache_path)
        ctx->autotune_path = options->autotune_cache_path;
      load_autotune_cache(ctx);
      DPAH_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
      ctx->d_counters = static_cast<unsigned long long*>(ctx->alloc(dpah::kCounterCount * sizeof(unsigned long long)));
      DPAH_HIP(hipMemsetAsync(ctx->d_counters, 0, dpah::kCounterCount * sizeof(unsigned long long), ctx->stream));
      ctx->sync();
    } catch (...) {
      delete ctx;
      throw;
    }
    *out = ctx;
  });
}

DPAH_EXPORT int dpah_destroy(dpah_context* ctx) {
  return guarded([&] {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->resolve_timings();
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    ctx->drain_deferred();
    for (auto& a : ctx->user_allocs) ctx->release(a.first, a.second);
    if (ctx->d_counters) ctx->release(ctx->d_counters, dpah::kCounterCount * sizeof(unsigned long long));
    ctx->trim_cache();
    for (auto& kv : ctx->block_size) (void)hipFree(kv.first);   // blocks never released
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
  });
}

DPAH_EXPORT int dpah_synchronize(dpah_context* ctx) {
  return guarded([&] {
    DPAH_REQUIRE(ctx, "ctx is null");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    ctx->sync();
  });
}

DPAH_EXPORT int dpah_device_alloc(dpah_context* ctx, uint64_t bytes, void** out) {
  return guarded([&] {
    DPAH_REQUIRE(ctx && out, "null argument");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    *out = ctx->alloc(bytes);
    ctx->user_allocs[*out] = bytes;
  });
}

DPAH_EXPORT int dpah_device_free(dpah_context* ctx, void* ptr) {
  return guarded([&] {
    DPAH_REQUIRE(ctx, "ctx is null");
    if (!ptr) return;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    DPAH_HIP(hipStreamSynchronize(ctx->stream));
    auto it = ctx->user_allocs.find(ptr);
    DPAH_REQUIRE(it != ctx->user_allocs.end(), "pointer was not returned by dpah_device_alloc");
    ctx->release(ptr, it->second);
    ctx->user_allocs.erase(it);
  });
}

DPAH_EXPORT int dpah_copy_to_device(dpah_context* ctx, void* dst, const void* src, uint64_t bytes) {
  return guarded([&] {
    DPAH_REQUIRE(ctx && (bytes == 0 || (dst && src)), "null argument");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    if (bytes) DPAH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    ctx->sync();
  });
}

DPAH_EXPORT int dpah_copy_to_host(dpah_context* ctx, void* dst, const void* src, uint64_t bytes) {
  return guarded([&] {
    DPAH_REQUIRE(ctx && (bytes == 0 || (dst && src)), "null argument");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    ctx->download(dst, src, bytes, ctx->stream);
    ctx->sync();
  });
}

// -------------------------------------------------------------- fields ----

// One thread per cell record: finds the record's cell, gathers its eight
// corners from the node array and writes them, in corner order 4a+2b+c, as one
// 128-B record.  Padding records of a brick-ordered layout (cells beyond the
// grid) are left unwritten: no sample inside the grid reads them.
static __global__ __launch_bounds__(256) void k_build_cells(FieldView v, float4* out,
                                                            uint64_t records) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= records) return;
  const uint64_t m2 = (uint64_t)v.n2 - 1, m1 = (uint64_t)v.n1 - 1, m0 = (uint64_t)v.n0 - 1;
  uint64_t c0, c1, c2, it;
  if (v.cell_bits) {
    const uint32_t k = v.cell_bits;
    const uint64_t m = (1ull << k) - 1;
    it = r / (uint64_t)v.cell_slice;
    const uint64_t rr = r - it * (uint64_t)v.cell_slice;
    const uint64_t brick = rr >> (3 * k), inner = rr & ((1ull << (3 * k)) - 1);
    const uint64_t b2 = brick % (uint64_t)v.cb2, b1 = brick / (uint64_t)v.cb2 % (uint64_t)v.cb1;
    const uint64_t b0 = brick / (uint64_t)v.cb2 / (uint64_t)v.cb1;
    c0 = (b0 << k) | (inner >> (2 * k));
    c1 = (b1 << k) | ((inner >> k) & m);
    c2 = (b2 << k) | (inner & m);
    if (c0 >= m0 || c1 >= m1 || c2 >= m2) return;
  } else {
    c2 = r % m2;
    const uint64_t r1 = r / m2;
    c1 = r1 % m1;
    const uint64_t r0 = r1 / m1;
    c0 = r0 % m0;
    it = r0 / m0;
  }
  const uint64_t n2 = (uint64_t)v.n2, n12 = (uint64_t)v.n1 * v.n2;
  const float* s = v.data + 3 * it * (uint64_t)v.slice;
  const uint64_t b = c0 * n12 + c1 * n2 + c2;
  const uint64_t off[8] = {b, b + 1, b + n2, b + n2 + 1, b + n12, b + n12 + 1, b + n12 + n2,
                           b + n12 + n2 + 1};
  float c[32];
  for (int k = 0; k < 8; ++k) {
    const float* p = s + 3 * off[k];
    c[3 * k] = p[0]; c[3 * k + 1] = p[1]; c[3 * k + 2] = p[2];
  }
  for (int k = 24; k < 32; ++k) c[k] = 0.0f;
  float4* o = out + 8 * r;
  for (int k = 0; k < 8; ++k) o[k] = make_float4(c[4 * k], c[4 * k + 1], c[4 * k + 2], c[4 * k + 3]);
}

// Mismatches of div_by(x, h, r) against x / h over every float x in [-B, B].
static uint64_t verify_division(dpah_context* ctx, float h, float r, float B) {
  uint32_t hi;
  std::memcpy(&hi, &B, 4);
  unsigned long long* bad = ctx->d_counters + 3;
  DPAH_HIP(hipMemsetAsync(bad, 0, sizeof(unsigned long long), ctx->stream));
  const unsigned grid = (unsigned)ctx->cu_count * 16;
  hipLaunchKernelGGL(verify_division_kernel, dim3(grid), dim3(256), 0, ctx->stream, h, r, hi, bad);
  DPAH_HIP(hipGetLastError());
  unsigned long long n = 0;
  DPAH_HIP(hipMemcpyAsync(&n, bad, sizeof(n), hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  return n;
}

// Reciprocal spacings and, per axis, whether the reciprocal division matches
// the correctly rounded one over the axis's whole range (checked once per
// (spacing, range) on a context).
static void set_division(dpah_context* ctx, FieldView& v) {
  const float h[4] = {v.h0, v.h1, v.h2, v.th};
  const int32_t n[4] = {v.n0, v.n1, v.n2, v.nt};
  float r[4];
  v.fastdiv = 0;
  v.exactdiv = 0;
  for (int a = 0; a < 4; ++a) {
    r[a] = 1.0f / h[a];
    if (n[a] < 2 || !ctx->tune.fastdiv) continue;
    int e = 0;
    if (std::frexp(h[a], &e) == 0.5f && std::isnormal(r[a])) {   // h = 2^(e-1), r exact
      v.exactdiv |= 1u << a;
      continue;
    }
    // the fast remainder's division needs non-zero remainders of at least
    // 2^-90: a spacing's ulp is one when the spacing is at least 2^-60
    if (!(h[a] >= 0x1p-60f)) continue;
    const float B = std::nextafter(2.0f * (float)(n[a] - 1) * h[a], INFINITY);
    if (!(B < INFINITY)) continue;
    uint32_t hb, bb;
    std::memcpy(&hb, &h[a], 4);
    std::memcpy(&bb, &B, 4);
    const uint64_t key = ((uint64_t)hb << 32) | bb;
    auto it = ctx->division_checked.find(key);
    bool ok;
    if (it != ctx->division_checked.end()) {
      ok = it->second;
    } else {
      ok = verify_division(ctx, h[a], r[a], B) == 0;
      ctx->division_checked[key] = ok;
    }
    if (ok) v.fastdiv |= 1u << a;
  }
  v.r0 = r[0]; v.r1 = r[1]; v.r2 = r[2]; v.rt = r[3];
}

// Builds the cell layout of a field when the context asks for it and the copy
// (128 B per cell and time slice) fits a tenth of the memory budget.
static void build_cells(dpah_context* ctx, dpah_field* f) {
  FieldView& v = f->view;
  const uint32_t k = (uint32_t)ctx->tune.cell_brick;
  uint64_t per_slice;
  if (k) {   // the cell grid padded to whole bricks of 2^k cells per axis
    auto bricks = [&](int32_t n) { return ((uint64_t)n - 1 + (1ull << k) - 1) >> k; };
    v.cb1 = (int32_t)bricks(v.n1);
    v.cb2 = (int32_t)bricks(v.n2);
    per_slice = (bricks(v.n0) * bricks(v.n1) * bricks(v.n2)) << (3 * k);
  } else {
    per_slice = (uint64_t)(v.n0 - 1) * (v.n1 - 1) * (v.n2 - 1);
  }
  const uint64_t records = per_slice * (uint64_t)v.nt;
  const uint64_t bytes = records * 128;
  if (ctx->tune.cells == 0 || bytes > ctx->memory_budget / 10) return;
  v.cell_bits = k;
  v.cell_slice = (int64_t)per_slice;
  f->d_cells = static_cast<float4*>(ctx->alloc(bytes));
  f->cell_bytes = bytes;
  const uint64_t grid = (records + 255) / 256;
  DPAH_REQUIRE(grid < (1ull << 31), "field too large for the cell layout");
  hipLaunchKernelGGL(k_build_cells, dim3((unsigned)grid), dim3(256), 0, ctx->stream, v,
                     f->d_cells, records);
  DPAH_HIP(hipGetLastError());
  ctx->sync();
  v.cells = f->d_cells;
  if (8 * records >= (1ull << 32)) v.idx32 = 0;
  // buffer reloads: the layout below 2^32 bytes and the factors of
  // cell_byte24's products (record counts per axis and plane) below 2^24
  const uint64_t plane = k ? (uint64_t)v.cb1 * (uint64_t)v.cb2
                           : (uint64_t)(v.n1 - 1) * (uint64_t)(v.n2 - 1);
  v.cell_bytes32 = bytes < (1ull << 32) && plane < (1ull << 24) ? (uint32_t)bytes : 0u;
}

// One thread per node of every slice: the node's 12 bytes to its place in the
// node bricks (FieldView::nbrick).
static __global__ __launch_bounds__(256) void k_build_nbrick(FieldView v, float* out,
                                                             uint64_t total) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const uint64_t t = i / (uint64_t)v.slice, r = i - t * (uint64_t)v.slice;
  const uint32_t n12 = (uint32_t)v.n1 * (uint32_t)v.n2;
  const uint32_t i0 = (uint32_t)(r / n12), q = (uint32_t)(r - (uint64_t)i0 * n12);
  const uint32_t i1 = q / (uint32_t)v.n2, i2 = q - i1 * (uint32_t)v.n2;
  const uint64_t brick = ((uint64_t)(i0 >> 2) * v.nb1 + (i1 >> 2)) * v.nb2 + (i2 >> 2);
  const uint64_t local = ((i0 & 3u) << 4) | ((i1 & 3u) << 2) | (i2 & 3u);
  float* o = (float*)((char*)out + t * v.nbrick_slice + brick * 768 + local * 12);
  const float* p = v.data + 3 * i;
  o[0] = p[0]; o[1] = p[1]; o[2] = p[2];
}

// The node bricks of a field, when the tuning asks for them (node_brick 1:
// steady fields, 2: every field) and they fit: every slice's bricks below
// 2^32 bytes in all and a slice's bricks numbering below 2^24 (the 24-bit
// offsets of nbrick_byte24), a tenth of the memory budget at most.
static void build_nbrick(dpah_context* ctx, dpah_field* f) {
  FieldView& v = f->view;
  if (ctx->tune.node_brick == 0 || (v.nt > 1 && ctx->tune.node_brick < 2)) return;
  const uint64_t nb0 = ((uint64_t)v.n0 + 3) / 4, nb1 = ((uint64_t)v.n1 + 3) / 4,
                 nb2 = ((uint64_t)v.n2 + 3) / 4;
  const uint64_t per_slice = nb0 * nb1 * nb2 * 768;
  const uint64_t bytes = per_slice * (uint64_t)v.nt;
  if (nb0 * nb1 * nb2 >= (1ull << 24) || bytes >= (1ull << 32) ||
      bytes > ctx->memory_budget / 10)
    return;
  v.nb1 = (uint32_t)nb1;
  v.nb2 = (uint32_t)nb2;
  v.nbrick_slice = (uint32_t)per_slice;
  f->d_nbrick = static_cast<float*>(ctx->alloc(bytes));
  f->nbrick_bytes = bytes;
  const uint64_t total = (uint64_t)v.slice * (uint64_t)v.nt;
  hipLaunchKernelGGL(k_build_nbrick, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     ctx->stream, v, f->d_nbrick, total);
  DPAH_HIP(hipGetLastError());
  ctx->sync();
  v.nbrick = (const char*)f->d_nbrick;
}

static void make_field(dpah_context* ctx, const float* xyzv, const uint64_t shape[3], uint64_t nt,
                       const float offset[3], const float spacing[3], float t_off, float t_sp,
                       dpah_field** out) {
  DPAH_REQUIRE(ctx && xyzv && shape && offset && spacing && out, "null argument");
  for (int i = 0; i < 3; ++i) {
    DPAH_REQUIRE(shape[i] >= 2 && shape[i] < (1ull << 24), "field shape must be in [2, 2^24) per axis");
    DPAH_REQUIRE(spacing[i] > 0.0f, "field spacing must be positive");
  }
  DPAH_REQUIRE(nt >= 1 && nt < (1ull << 24), "time_steps must be in [1, 2^24)");
  if (nt > 1) DPAH_REQUIRE(nt >= 2 && t_sp > 0.0f, "time-variant field needs time_spacing > 0");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  DPAH_HIP(hipSetDevice(ctx->device));
  const uint64_t nodes = shape[0] * shape[1] * shape[2];
  const uint64_t bytes = nodes * nt * 3 * sizeof(float);
  auto* f = new dpah_field();
  try {
    f->ctx = ctx;
    f->bytes = bytes;
    f->d_data = static_cast<float*>(ctx->alloc(bytes));
    DPAH_HIP(hipMemcpyAsync(f->d_data, xyzv, bytes, hipMemcpyHostToDevice, ctx->stream));
    ctx->sync();
  } catch (...) {
    if (f->d_data) ctx->release(f->d_data, bytes);
    delete f;
    throw;
  }
  for (int i = 0; i < 3; ++i) f->shape[i] = shape[i];
  f->time_steps = nt;
  FieldView& v = f->view;
  v.data = f->d_data;
  v.slice = (int64_t)nodes;
  v.n0 = (int32_t)shape[0]; v.n1 = (int32_t)shape[1]; v.n2 = (int32_t)shape[2];
  v.nt = (int32_t)nt;
  v.o0 = offset[0]; v.o1 = offset[1]; v.o2 = offset[2]; v.to = t_off;
  v.h0 = spacing[0]; v.h1 = spacing[1]; v.h2 = spacing[2]; v.th = nt > 1 ? t_sp : 1.0f;
  v.cells = nullptr;
  v.cell_slice = 0;
  v.cell_bits = 0;
  v.cb1 = v.cb2 = 0;
  v.idx32 = 3 * nodes * nt < (1ull << 32) ? 1u : 0u;
  // buffer reloads: the node array below 2^32 bytes and a plane's bytes
  // 12 n1 n2 below 2^24 (every axis already is), the factors of node_byte24's
  // products
  v.node_bytes32 =
      bytes < (1ull << 32) && 12 * shape[1] * shape[2] < (1ull << 24) ? (uint32_t)bytes : 0u;
  v.cell_bytes32 = 0;
  v.nbrick = nullptr;
  v.nb1 = v.nb2 = v.nbrick_slice = 0;
  try {
    set_division(ctx, v);
    build_cells(ctx, f);
    build_nbrick(ctx, f);
  } catch (...) {
    if (f->d_nbrick) ctx->release(f->d_nbrick, f->nbrick_bytes);
    if (f->d_cells) ctx->release(f->d_cells, f->cell_bytes);
    ctx->release(f->d_data, f->bytes);
    delete f;
    throw;
  }
  *out = f;
}

DPAH_EXPORT int dpah_field_create(dpah_context* ctx, const float* xyzv, const uint64_t shape[3],
                                  const float offset[3], const float spacing[3], dpah_field** out) {
  return guarded([&] { make_field(ctx, xyzv, shape, 1, offset, spacing, 0.0f, 1.0f, out); });
}

DPAH_EXPORT int dpah_field_create_unsteady(dpah_context* ctx, const float* xyzv,
                                           const uint64_t shape[3], uint64_t time_steps,
                                           const float offset[3], const float spacing[3],
                                           float time_offset, float time_spacing,
                                           dpah_field** out) {
  return guarded([&] {
    DPAH_REQUIRE(time_steps >= 2, "a time-variant field needs at least two time steps");
    make_field(ctx, xyzv, shape, time_steps, offset, spacing, time_offset, time_spacing, out);
  });
}

DPAH_EXPORT int dpah_field_release(dpah_field* field) {
  return guarded([&] {
    if (!field) return;
    dpah_context* ctx = field->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    DPAH_HIP(hipStreamSynchronize(ctx->stream));
    ctx->release(field->d_data, field->bytes);
    if (field->d_cells) ctx->release(field->d_cells, field->cell_bytes);
    if (field->d_nbrick) ctx->release(field->d_nbrick, field->nbrick_bytes);
    delete field;
  });
}

// ------------------------------------------------------------- advect ----

namespace {

constexpr uint64_t kMaxRemaining = (1ull << 32) - 3;

void check_integrator(int integrator) {
  DPAH_REQUIRE(integrator == DPAH_INTEGRATOR_DPA_RK4 || integrator == DPAH_INTEGRATOR_RK4,
               "unknown integrator");
}

}  // namespace

dpah_result* dpah::run_advect(dpah_context* ctx, const dpah_field* field, const float* seeds,
                              const uint64_t* remaining, uint64_t uniform_remaining,
                              uint64_t count, const dpah_advect_params* params, bool persistent) {
  {
    DPAH_REQUIRE(ctx && field && params, "null argument");
    DPAH_REQUIRE(field->ctx == ctx, "field belongs to another context");
    DPAH_REQUIRE(count == 0 || seeds, "seeds is null");
    DPAH_REQUIRE(count == 0 || params->particles_per_round > 0,
                 "particles_per_round must be positive (dpa would never finish)");
    check_integrator(params->integrator);
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));

    std::vector<uint64_t> rem(count);
    for (uint64_t i = 0; i < count; ++i) {
      rem[i] = remaining ? remaining[i] : uniform_remaining;
      DPAH_REQUIRE(rem[i] <= kMaxRemaining, "remaining iterations must be below 2^32 - 2");
    }

    auto res = std::make_unique<dpah_result>();
    res->ctx = ctx;
    res->use64 = params->use_64_bit_indices ? 1 : 0;
    res->final_pos.resize(3 * count);
    res->final_rem.resize(count);
    res->final_status.resize(count);

    const uint64_t n1 = count ? count : 1;
    DeviceBuffer d_seeds(ctx, 3 * n1 * sizeof(float));
    DeviceBuffer d_rem(ctx, n1 * sizeof(uint64_t));
    DeviceBuffer d_pos(ctx, 3 * n1 * sizeof(float));
    DeviceBuffer d_rem_out(ctx, n1 * sizeof(uint64_t));
    DeviceBuffer d_status(ctx, n1);
    if (count) {
      DPAH_HIP(hipMemcpyAsync(d_seeds.ptr, seeds, 3 * count * sizeof(float), hipMemcpyHostToDevice,
                              ctx->stream));
      DPAH_HIP(hipMemcpyAsync(d_rem.ptr, rem.data(), count * sizeof(uint64_t),
                              hipMemcpyHostToDevice, ctx->stream));
    }

    AdvectLaunch L{};
    L.f = &field->view;
    L.uniform_rem = 0;
    L.dt = params->step_size;
    L.t0 = params->start_time;
    L.integrator = params->integrator;
    L.persistent = persistent;

    uint64_t total_steps = 0;
    auto read_steps = [&] {
      unsigned long long s = 0;
      DPAH_HIP(hipMemcpyAsync(&s, ctx->d_counters + 1, sizeof(s), hipMemcpyDeviceToHost,
                              ctx->stream));
      ctx->sync();
      total_steps += s;
    };

    if (!params->record) {
      // Rounds only decide the order in which independent particles run, and
      // with record off they produce no output: one launch covers them all.
      L.seeds = d_seeds.as<float>();
      L.rem_in = d_rem.as<uint64_t>();
      L.pos_out = d_pos.as<float>();
      L.rem_out = d_rem_out.as<uint64_t>();
      L.status_out = d_status.as<uint8_t>();
      L.n = count;
      launch_advect(ctx, L);
      read_steps();
    } else {
      // a single request's recorded rounds run one thread per row unless the
      // tuning asks for the persistent kernel (record_grid)
      if (persistent && record_grid(ctx, field->view, params->integrator, false))
        persistent = L.persistent = false;
      const uint64_t ppr = params->particles_per_round;
      DeviceBuffer d_steps(ctx, n1 * sizeof(uint32_t));
      DeviceBuffer d_special(ctx, n1);
      uint64_t active = count;
      do {  // pipeline.cpp:100-146: at least one round, tail of the seeds first
        const uint64_t rc = std::min(ppr, active);
        const uint64_t first = active - rc;
        uint64_t max_rem = 0;
        for (uint64_t i = 0; i < rc; ++i) max_rem = std::max(max_rem, rem[first + i]);
        const uint64_t stride = max_rem + 2;   // compute_round_state (305)
        Round round;
        if (rc) {
          // chunks of whole 64-row tiles (the slot layout), within the slot budget
          // (sized for the larger, staged block layout)
          const uint64_t row_bytes = slot_region_floats(1, stride, kStagedShift) / kTileRows *
                                     sizeof(float);
          uint64_t chunk = std::max<uint64_t>(1, ctx->slot_budget / row_bytes) / kTileRows *
                           kTileRows;
          chunk = std::min(std::max<uint64_t>(chunk, kTileRows), rc);
          const uint32_t g = record_slot_shift(ctx, field->view, params->integrator, chunk,
                                               persistent);
          DeviceBuffer slots(ctx, slot_region_floats(chunk, stride, g) * sizeof(float));
          for (uint64_t c0 = 0; c0 < rc; c0 += chunk) {
            const uint64_t cn = std::min(chunk, rc - c0);
            const uint64_t pid0 = first + c0;
            L.seeds = d_seeds.as<float>() + 3 * pid0;
            L.rem_in = d_rem.as<uint64_t>() + pid0;
            L.pos_out = d_pos.as<float>() + 3 * pid0;
            L.rem_out = d_rem_out.as<uint64_t>() + pid0;
            L.status_out = d_status.as<uint8_t>() + pid0;
            L.steps_out = d_steps.as<uint32_t>() + pid0;
            L.special_out = d_special.as<uint8_t>() + pid0;
            L.slots = slots.as<float>();
            L.stride = stride;
            L.slot_shift = g;
            L.n = cn;
            launch_advect(ctx, L);
            read_steps();
            Chunk ch = compact_rows(ctx, slots.as<float>(), stride, g, L.steps_out,
                                    L.special_out, cn, round.V, res->use64);
            round.V += ch.V;
            round.I += ch.I;
            round.chunks.push_back(ch);
          }
          if (!res->use64)
            DPAH_REQUIRE(round.V <= 0xffffffffull,
                         "round has more than 2^32-1 vertices; 32-bit indices cannot address it");
        }
        res->rounds.push_back(std::move(round));
        active = first;
      } while (active > 0);
    }
    uint8_t* stage = count ? static_cast<uint8_t*>(ctx->host_staging(21 * count)) : nullptr;
    if (count) {
      // final states through the pinned staging when they fit
      void* pos = stage ? static_cast<void*>(stage + 8 * count) : res->final_pos.data();
      void* rem = stage ? static_cast<void*>(stage) : res->final_rem.data();
      void* st = stage ? static_cast<void*>(stage + 20 * count) : res->final_status.data();
      DPAH_HIP(hipMemcpyAsync(pos, d_pos.ptr, 3 * count * sizeof(float), hipMemcpyDeviceToHost,
                              ctx->stream));
      DPAH_HIP(hipMemcpyAsync(rem, d_rem_out.ptr, count * sizeof(uint64_t),
                              hipMemcpyDeviceToHost, ctx->stream));
      DPAH_HIP(hipMemcpyAsync(st, d_status.ptr, count, hipMemcpyDeviceToHost, ctx->stream));
    }
    ctx->sync();
    if (stage) {
      std::memcpy(res->final_rem.data(), stage, 8 * count);
      std::memcpy(res->final_pos.data(), stage + 8 * count, 12 * count);
      std::memcpy(res->final_status.data(), stage + 20 * count, count);
    }
    res->total_steps = total_steps;
    return res.release();
  }
}

DPAH_EXPORT int dpah_advect(dpah_context* ctx, const dpah_field* field, const float* seeds,
                            const uint64_t* remaining, uint64_t uniform_remaining, uint64_t count,
                            const dpah_advect_params* params, dpah_result** out) {
  return guarded([&] {
    DPAH_REQUIRE(out, "null argument");
    *out = run_advect(ctx, field, seeds, remaining, uniform_remaining, count, params, true);
  });
}

DPAH_EXPORT int dpah_result_round_count(const dpah_result* r, uint64_t* rounds) {
  return guarded([&] {
    DPAH_REQUIRE(r && rounds, "null argument");
    *rounds = r->rounds.size();
  });
}

DPAH_EXPORT int dpah_result_round_sizes(const dpah_result* r, uint64_t round, uint64_t* vertex_count,
                                        uint64_t* index_count, int* index_bytes) {
  return guarded([&] {
    DPAH_REQUIRE(r, "null argument");
    DPAH_REQUIRE(round < r->rounds.size(), "round out of range");
    if (vertex_count) *vertex_count = r->rounds[round].V;
    if (index_count) *index_count = r->rounds[round].I;
    if (index_bytes) *index_bytes = r->use64 ? 8 : 4;
  });
}

DPAH_EXPORT int dpah_result_copy_round(const dpah_result* r, uint64_t round, float* vertices,
                                       float* colors, void* indices) {
  return guarded([&] {
    DPAH_REQUIRE(r, "null argument");
    DPAH_REQUIRE(round < r->rounds.size(), "round out of range");
    dpah_context* ctx = r->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    const uint64_t ib = r->use64 ? 8 : 4;
    uint64_t v0 = 0, i0 = 0;
    for (const Chunk& c : r->rounds[round].chunks) {
      if (!c.vertices) {   // materialised on the host
        const uint8_t* h = c.host.get();
        if (vertices && c.V) std::memcpy(vertices + 3 * v0, h, 12 * c.V);
        if (colors && c.V) std::memcpy(colors + v0, h + c.colors_off, 4 * c.V);
        if (indices && c.I)
          std::memcpy(static_cast<char*>(indices) + i0 * ib, h + c.indices_off, c.I * ib);
        v0 += c.V;
        i0 += c.I;
        continue;
      }
      if (vertices && c.V)
        ctx->download(vertices + 3 * v0, c.vertices, 3 * c.V * sizeof(float), ctx->stream);
      if (colors && c.V) ctx->download(colors + v0, c.colors, c.V * sizeof(float), ctx->stream);
      if (indices && c.I)
        ctx->download(static_cast<char*>(indices) + i0 * ib, c.indices, c.I * ib, ctx->stream);
      v0 += c.V;
      i0 += c.I;
    }
    ctx->sync();
  });
}

DPAH_EXPORT int dpah_result_final_state(const dpah_result* r, float* positions, uint64_t* remaining,
                                        uint8_t* status) {
  return guarded([&] {
    DPAH_REQUIRE(r, "null argument");
    const uint64_t n = r->final_rem.size();
    if (positions && n) std::memcpy(positions, r->final_pos.data(), 3 * n * sizeof(float));
    if (remaining && n) std::memcpy(remaining, r->final_rem.data(), n * sizeof(uint64_t));
    if (status && n) std::memcpy(status, r->final_status.data(), n);
  });
}

DPAH_EXPORT int dpah_result_total_steps(const dpah_result* r, uint64_t* steps) {
  return guarded([&] {
    DPAH_REQUIRE(r && steps, "null argument");
    *steps = r->total_steps;
  });
}

DPAH_EXPORT int dpah_result_free(dpah_result* r) {
  return guarded([&] {
    if (!r) return;
    dpah_context* ctx = r->ctx;
    std::unique_lock<std::recursive_mutex> lk(ctx->mu, std::try_to_lock);
    if (!lk.owns_lock()) {   // another thread's call holds the context: defer
      std::lock_guard<std::mutex> d(ctx->deferred_mu);
      ctx->deferred.push_back(r);
      return;
    }
    DPAH_HIP(hipSetDevice(ctx->device));
    delete r;
  });
}

DPAH_EXPORT int dpah_advect_endpoints_device(dpah_context* ctx, const dpah_field* field,
                                             const float* seeds, const uint64_t* remaining,
                                             uint64_t uniform_remaining, uint64_t count,
                                             float step_size, float start_time, int integrator,
                                             float* positions_out, uint64_t* remaining_out,
                                             uint8_t* status_out, uint64_t* total_steps) {
  return guarded([&] {
    DPAH_REQUIRE(ctx && field, "null argument");
    DPAH_REQUIRE(field->ctx == ctx, "field belongs to another context");
    DPAH_REQUIRE(count == 0 || (seeds && positions_out), "seeds / positions_out is null");
    DPAH_REQUIRE(remaining || uniform_remaining <= kMaxRemaining,
                 "remaining iterations must be below 2^32 - 2");
    check_integrator(integrator);
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
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
    launch_advect(ctx, L);
    if (total_steps) {
      unsigned long long s = 0;
      DPAH_HIP(hipMemcpyAsync(&s, ctx->d_counters + 1, sizeof(s), hipMemcpyDeviceToHost,
                              ctx->stream));
      ctx->sync();
      *total_steps = s;
    }
  });
}

DPAH_EXPORT int dpah_kernel_timing(dpah_context* ctx, const char* kernel, uint64_t* launches,
                                   double* total_ms) {
  return guarded([&] {
    DPAH_REQUIRE(ctx && kernel, "null argument");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    ctx->sync();
    auto it = ctx->timings.find(kernel);
    if (launches) *launches = it == ctx->timings.end() ? 0 : it->second.launches;
    if (total_ms) *total_ms = it == ctx->timings.end() ? 0.0 : it->second.total_ms;
  });
}

DPAH_EXPORT int dpah_set_count_cell_loads(dpah_context* ctx, int on) {
  return guarded([&] {
    DPAH_REQUIRE(ctx, "ctx is null");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    if (on && !ctx->count_loads) dpah::clear_load_counters(ctx);
    ctx->count_loads = on != 0;
  });
}

DPAH_EXPORT int dpah_cell_loads(dpah_context* ctx, uint64_t* record_loads, uint64_t* node_loads) {
  return guarded([&] {
    DPAH_REQUIRE(ctx, "ctx is null");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    unsigned long long c[2] = {0, 0};
    DPAH_HIP(hipMemcpyAsync(c, ctx->d_counters + dpah::kLoadsCells, sizeof(c),
                            hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    if (record_loads) *record_loads = c[0];
    if (node_loads) *node_loads = c[1];
  });
}

DPAH_EXPORT int dpah_kernel_timing_reset(dpah_context* ctx) {
  return guarded([&] {
    DPAH_REQUIRE(ctx, "null argument");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    ctx->sync();
    ctx->timings.clear();
  });
}

DPAH_EXPORT int dpah_get_tuning(dpah_context* ctx, dpah_tuning* out) {
  return guarded([&] {
    DPAH_REQUIRE(ctx && out, "null argument");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    out->block = ctx->tune.block;
    out->blocks_per_cu = ctx->tune.blocks_per_cu;
    out->order = ctx->tune.order;
    out->parts = ctx->tune.parts;
    out->refill = ctx->tune.refill;
    out->cells = ctx->tune.cells;
    out->fastdiv = ctx->tune.fastdiv;
    out->cell_brick = ctx->tune.cell_brick;
    out->order_brick = ctx->tune.order_brick;
    out->lattice_brick = ctx->tune.lattice_brick;
    out->slot_stage = ctx->tune.slot_stage;
    out->lattice_group = ctx->tune.lattice_group;
    out->node_brick = ctx->tune.node_brick;
  });
}

DPAH_EXPORT int dpah_set_tuning(dpah_context* ctx, const dpah_tuning* t) {
  return guarded([&] {
    DPAH_REQUIRE(ctx && t, "null argument");
    DPAH_REQUIRE(t->block >= 64 && t->block <= 256 && t->block % 64 == 0,
                 "block must be 64, 128, 192 or 256");
    DPAH_REQUIRE(t->blocks_per_cu >= 0 && t->blocks_per_cu <= 32,
                 "blocks_per_cu must be 0 (by launch kind) or 1..32");
    DPAH_REQUIRE(t->order == 0 || t->order == 1, "order must be 0 or 1");
    DPAH_REQUIRE(t->parts >= 1 && t->parts <= dpah::kMaxParts, "parts must be 1..8");
    DPAH_REQUIRE(t->refill >= 0 && t->refill <= 64, "refill must be 0 (auto) or 1..64");
    DPAH_REQUIRE(t->cells >= 0 && t->cells <= 2, "cells must be 0, 1 or 2");
    DPAH_REQUIRE(t->fastdiv == 0 || t->fastdiv == 1, "fastdiv must be 0 or 1");
    DPAH_REQUIRE(t->cell_brick == 0 || (t->cell_brick >= 2 && t->cell_brick <= 5),
                 "cell_brick must be 0 or 2..5");
    DPAH_REQUIRE(t->order_brick >= 0 && t->order_brick <= 5, "order_brick must be 0..5");
    DPAH_REQUIRE(t->lattice_brick == 0 || t->lattice_brick == 2 || t->lattice_brick == 4 ||
                     t->lattice_brick == 8,
                 "lattice_brick must be 0, 2, 4 or 8");
    DPAH_REQUIRE(t->slot_stage >= 0 && t->slot_stage <= 3, "slot_stage must be 0, 1, 2 or 3");
    DPAH_REQUIRE(t->lattice_group == 0 || t->lattice_group == 1 || t->lattice_group == 2 ||
                     t->lattice_group == 4 || t->lattice_group == 8 || t->lattice_group == 16,
                 "lattice_group must be 0, 1, 2, 4, 8 or 16");
    DPAH_REQUIRE(t->node_brick >= 0 && t->node_brick <= 2, "node_brick must be 0, 1 or 2");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    ctx->tune.block = t->block;
    ctx->tune.blocks_per_cu = t->blocks_per_cu;
    ctx->tune.order = t->order;
    ctx->tune.parts = t->parts;
    ctx->tune.refill = t->refill;
    ctx->tune.cells = t->cells;
    ctx->tune.fastdiv = t->fastdiv;
    ctx->tune.cell_brick = t->cell_brick;
    ctx->tune.order_brick = t->order_brick;
    ctx->tune.lattice_brick = t->lattice_brick;
    ctx->tune.slot_stage = t->slot_stage;
    ctx->tune.lattice_group = t->lattice_group;
    ctx->tune.node_brick = t->node_brick;
    ctx->tuned.clear();   // explicit settings replace autotuned ones
  });
}

DPAH_EXPORT int dpah_verify_division(dpah_context* ctx, float h, float r, float bound,
                                     uint64_t* mismatches) {
  return guarded([&] {
    DPAH_REQUIRE(ctx && mismatches, "null argument");
    DPAH_REQUIRE(h > 0.0f && bound >= 0.0f && bound < INFINITY, "need h > 0 and a finite bound");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    *mismatches = verify_division(ctx, h, r, bound);
  });
}
