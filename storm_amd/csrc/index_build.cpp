// index_build.cpp -- genome planes and seed indexes built on the device (load_reference_sequence's
// buffers, reference_data.c:14-190, and index__build_reference, index.c:19-195), and the C ABI
// of the standalone index build (storm_index_*).
#include <cstring>

#include "mapper.h"

namespace storm {

namespace {

int64_t plane_words(int64_t size) { return (GPAD + size + GPAD) / 32 + 4; }

SeedDev seed_dev(const SeedSpec& sp) {
  SeedDev s{};
  s.weight = sp.weight;
  s.span = sp.span;
  s.n_pos = 0;
  for (int j = 0; j < sp.weight; ++j) s.offs[j] = sp.offs[j];
  return s;
}

}  // namespace

IndexWork::~IndexWork() {
  if (host_counts) (void)hipHostFree(host_counts);
}

int upload_genome(const uint8_t* code, const uint8_t* mask, int64_t size, DevGenome& g, hipStream_t st,
                  DevMem* keep_code) {
  const int64_t nw = plane_words(size);
  DevMem dcode, dmask;
  HIPOK(dcode.ensure((size_t)size + 64));
  if (size > 0) HIPOK(hipMemcpyAsync(dcode.p, code, (size_t)size, hipMemcpyHostToDevice, st));
  if (mask && size > 0) {
    HIPOK(dmask.ensure((size_t)size));
    HIPOK(hipMemcpyAsync(dmask.p, mask, (size_t)size, hipMemcpyHostToDevice, st));
  }
  HIPOK(g.lh.ensure(nw * 8));
  HIPOK(g.mask.ensure(nw * 4));
  HIPOK(launch_genome_planes(dcode.as<uint8_t>(), mask ? dmask.as<uint8_t>() : nullptr, size, nw, g.lh.as<uint2>(),
                             g.mask.as<uint32_t>(), st));
  HIPOK(hipStreamSynchronize(st));
  g.size = size;
  if (keep_code) *keep_code = std::move(dcode);
  return STORM_OK;
}

int build_index_dev(const uint8_t* d_code, int64_t size, const int2* d_regions, int n_regions, const SeedSpec& seed,
                    int stdev_distance, DevIndex& ix, IndexWork& w, hipStream_t st, int64_t* erased_out) {
  const int64_t ks = 1LL << (2 * seed.weight);
  const int64_t map_size = size - seed.span + 1;
  ix.spec = seed;
  if (erased_out) *erased_out = 0;
  HIPOK(ix.first.ensure((size_t)(ks + 1) * 4));
  if (map_size <= 0 || seed.span >= size) {   // index.c:26-31,64: nothing indexed
    HIPOK(hipMemsetAsync(ix.first.p, 0, (size_t)(ks + 1) * 4, st));
    HIPOK(ix.map.ensure(16));
    ix.n_map = 0;
    return STORM_OK;
  }
  const int64_t n = map_size;
  HIPOK(w.keys0.ensure((size_t)n * 4));
  HIPOK(w.keys1.ensure((size_t)n * 4));
  HIPOK(w.vals0.ensure((size_t)n * 4));
  HIPOK(w.vals1.ensure((size_t)n * 4));
  HIPOK(w.st.ensure((size_t)(ks + 1) * 4));
  HIPOK(w.en.ensure((size_t)(ks + 1) * 4));
  HIPOK(launch_index_keys(d_code, n, d_regions, n_regions, seed_dev(seed), (uint32_t)ks, w.keys0.as<uint32_t>(),
                          w.vals0.as<int32_t>(), st));
  const int end_bit = 2 * seed.weight + 1;   // the skip key 4^w needs one bit more than the keys
  int sel = 0;
  size_t tb = 0;
  HIPOK(sort_key_pos_raw(w.keys0.as<uint32_t>(), w.keys1.as<uint32_t>(), w.vals0.as<int32_t>(), w.vals1.as<int32_t>(),
                         n, end_bit, nullptr, tb, sel, st));
  HIPOK(w.tmp.ensure(tb));
  HIPOK(sort_key_pos_raw(w.keys0.as<uint32_t>(), w.keys1.as<uint32_t>(), w.vals0.as<int32_t>(), w.vals1.as<int32_t>(),
                         n, end_bit, w.tmp.p, tb, sel, st));
  const uint32_t* keys = sel ? w.keys1.as<uint32_t>() : w.keys0.as<uint32_t>();
  const int32_t* vals = sel ? w.vals1.as<int32_t>() : w.vals0.as<int32_t>();
  HIPOK(hipMemsetAsync(w.st.p, 0, (size_t)(ks + 1) * 4, st));
  HIPOK(hipMemsetAsync(w.en.p, 0, (size_t)(ks + 1) * 4, st));
  HIPOK(launch_index_runs(keys, n, w.st.as<int32_t>(), w.en.as<int32_t>(), st));
  HIPOK(launch_index_counts(w.st.as<int32_t>(), w.en.as<int32_t>(), ks, st));   // en = counts
  const uint8_t* d_erased = nullptr;
  if (stdev_distance > 0) {   // index.c:117-156 on the host, in the reference's summation order
    if (w.host_counts_n < ks) {
      if (w.host_counts) (void)hipHostFree(w.host_counts);
      w.host_counts = nullptr;
      w.host_counts_n = 0;
      HIPOK(hipHostMalloc((void**)&w.host_counts, (size_t)ks * 4, hipHostMallocDefault));
      w.host_counts_n = ks;
    }
    HIPOK(hipMemcpyAsync(w.host_counts, w.en.p, (size_t)ks * 4, hipMemcpyDeviceToHost, st));
    HIPOK(hipStreamSynchronize(st));
    int64_t count = 0;
    for (int64_t k = 0; k < ks; ++k) count += w.host_counts[k];
    w.host_erased.assign((size_t)ks, 0);
    const int64_t ne = repeat_filter(w.host_counts, ks, count, stdev_distance, w.host_erased.data());
    if (erased_out) *erased_out = ne;
    if (ne > 0) {
      HIPOK(w.erased.ensure((size_t)ks));
      HIPOK(hipMemcpyAsync(w.erased.p, w.host_erased.data(), (size_t)ks, hipMemcpyHostToDevice, st));
      HIPOK(launch_index_erase(w.erased.as<uint8_t>(), w.en.as<int32_t>(), ks, st));
      d_erased = w.erased.as<uint8_t>();
    }
  }
  size_t sb = 0;
  HIPOK(scan_exclusive_i32(w.en.as<int32_t>(), ix.first.as<int32_t>(), ks + 1, nullptr, sb, st));
  if (sb > w.tmp.n) HIPOK(w.tmp.ensure(sb));
  HIPOK(scan_exclusive_i32(w.en.as<int32_t>(), ix.first.as<int32_t>(), ks + 1, w.tmp.p, sb, st));
  HIPOK(ix.map.ensure((size_t)(n + 1) * 4 + 16));
  HIPOK(launch_index_scatter(keys, vals, n, w.st.as<int32_t>(), ix.first.as<int32_t>(), d_erased, (uint32_t)ks,
                             ix.map.as<int32_t>(), st));
  int32_t total = 0;
  HIPOK(hipMemcpyAsync(&total, ix.first.as<int32_t>() + ks, 4, hipMemcpyDeviceToHost, st));
  HIPOK(hipStreamSynchronize(st));
  ix.n_map = total;
  return STORM_OK;
}

int build_indexes(const HostRef& ref, const DevMem& d_code, const std::vector<SeedSpec>& seeds, int stdev_distance,
                  std::vector<DevIndex>& out, hipStream_t st, std::vector<int64_t>* erased) {
  std::vector<int2> reg(ref.regions.size());
  for (size_t k = 0; k < reg.size(); ++k) reg[k] = int2{ref.regions[k].first, ref.regions[k].second};
  DevMem dreg;
  HIPOK(dreg.ensure(reg.size() * sizeof(int2) + 16));
  if (!reg.empty()) HIPOK(hipMemcpyAsync(dreg.p, reg.data(), reg.size() * sizeof(int2), hipMemcpyHostToDevice, st));
  out.clear();
  out.resize(seeds.size());
  IndexWork w;
  if (erased) erased->assign(seeds.size(), 0);
  for (size_t s = 0; s < seeds.size(); ++s) {
    int rc = build_index_dev(d_code.as<uint8_t>(), ref.size, dreg.as<int2>(), (int)reg.size(), seeds[s],
                             stdev_distance, out[s], w, st, erased ? &(*erased)[s] : nullptr);
    if (rc) return rc;
  }
  HIPOK(hipStreamSynchronize(st));
  return STORM_OK;
}

}  // namespace storm

using namespace storm;

struct storm_index {
  int device = 0;
  DevIndex ix;
  int64_t n_keys = 0, erased = 0;
  double ms = 0;
};

extern "C" {

int storm_index_build(int device, const uint8_t* packed, int64_t size, const int32_t* region_start,
                      const int32_t* region_end, int n_regions, const char* seed, int stdev_distance,
                      storm_index** out) {
  if (!out) return fail(STORM_ERR_INVALID, "storm_index_build: null output");
  *out = nullptr;
  if (size < 0 || size > 0x7fffffffLL || (size > 0 && !packed))
    return fail(STORM_ERR_INVALID, "storm_index_build: sequence of %lld codes", (long long)size);
  if (n_regions < 0 || (n_regions > 0 && (!region_start || !region_end)))
    return fail(STORM_ERR_INVALID, "storm_index_build: bad masked region list");
  for (int k = 0; k < n_regions; ++k) {
    if (region_start[k] < 0 || region_start[k] >= region_end[k] || region_end[k] > size ||
        (k > 0 && region_start[k] < region_end[k - 1]))
      return fail(STORM_ERR_INVALID, "storm_index_build: masked region %d [%d, %d) is not sorted, disjoint and inside "
                  "the sequence", k, region_start[k], region_end[k]);
  }
  std::vector<SeedSpec> seeds;
  std::string err;
  int rc = parse_seeds(seed, seeds, err);
  if (rc) return fail(rc, "%s", err.c_str());
  if (seeds.size() != 1 || !seeds[0].pos.empty())
    return fail(STORM_ERR_INVALID, "storm_index_build: give exactly one seed pattern without positions");
  HIPOK(hipSetDevice(device));
  std::unique_ptr<storm_index> ix(new storm_index());
  ix->device = device;
  hipStream_t st = nullptr;
  HIPOK(hipStreamCreate(&st));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{st};
  HostRef ref;
  ref.size = size;
  for (int k = 0; k < n_regions; ++k) ref.regions.emplace_back(region_start[k], region_end[k]);
  DevMem dpacked, dcode;
  HIPOK(dpacked.ensure((size_t)(size + 3) / 4 + 16));
  HIPOK(dcode.ensure((size_t)size + 64));
  hipEvent_t e0, e1;
  HIPOK(hipEventCreate(&e0));
  HIPOK(hipEventCreate(&e1));
  if (size > 0) HIPOK(hipMemcpyAsync(dpacked.p, packed, (size_t)(size + 3) / 4, hipMemcpyHostToDevice, st));
  HIPOK(hipEventRecord(e0, st));
  HIPOK(launch_unpack_codes(dpacked.as<uint8_t>(), size, dcode.as<uint8_t>(), st));
  std::vector<DevIndex> out1;
  std::vector<int64_t> erased;
  rc = build_indexes(ref, dcode, seeds, stdev_distance, out1, st, &erased);
  if (rc == STORM_OK) {
    HIPOK(hipEventRecord(e1, st));
    HIPOK(hipEventSynchronize(e1));
    float ms = 0;
    HIPOK(hipEventElapsedTime(&ms, e0, e1));
    ix->ms = ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  ix->ix = std::move(out1[0]);
  ix->n_keys = 1LL << (2 * seeds[0].weight);
  ix->erased = erased[0];
  *out = ix.release();
  return STORM_OK;
}

int storm_index_sizes(const storm_index* ix, int64_t* n_keys, int64_t* n_map, int64_t* n_erased, double* ms) {
  if (!ix) return fail(STORM_ERR_INVALID, "storm_index_sizes: null index");
  if (n_keys) *n_keys = ix->n_keys;
  if (n_map) *n_map = ix->ix.n_map;
  if (n_erased) *n_erased = ix->erased;
  if (ms) *ms = ix->ms;
  return STORM_OK;
}

int storm_index_fetch(const storm_index* ix, int32_t* first, int32_t* map) {
  if (!ix) return fail(STORM_ERR_INVALID, "storm_index_fetch: null index");
  HIPOK(hipSetDevice(ix->device));
  if (first) HIPOK(hipMemcpy(first, ix->ix.first.p, (size_t)(ix->n_keys + 1) * 4, hipMemcpyDeviceToHost));
  if (map && ix->ix.n_map > 0) HIPOK(hipMemcpy(map, ix->ix.map.p, (size_t)ix->ix.n_map * 4, hipMemcpyDeviceToHost));
  return STORM_OK;
}

void storm_index_destroy(storm_index* ix) {
  if (!ix) return;
  (void)hipSetDevice(ix->device);
  delete ix;
}

}  // extern "C"
