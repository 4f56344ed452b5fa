// Exercises the C++ host API (include/dtl/filter.hpp) on the GPU the way the
// reference's own tests use its classes (src/dtl/filter/filter_test.cpp:16-40).
// Prints "PASS <name>" per check and exits non-zero on the first failure.
#include <dtl/filter.hpp>

#include <cstdio>
#include <cstdlib>
#include <vector>

#define CHECK(cond)                                                              \
  do {                                                                           \
    if (!(cond)) {                                                               \
      std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #cond);                 \
      std::exit(1);                                                              \
    }                                                                            \
  } while (0)

template <class T>
static T* dev_copy(const std::vector<T>& v) {
  T* d = nullptr;
  CHECK(hipMalloc(&d, std::max<size_t>(v.size(), 1) * sizeof(T)) == hipSuccess);
  CHECK(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) == hipSuccess);
  return d;
}

static void basic_build_probe_test() {
  // filter_test.cpp:16-40: default config (k=8, 32-bit words, w=1, s=1), m=1024
  dtl::blocked_bloomfilter_config config;
  dtl::filter::filter f(config, 1024);
  std::vector<uint32_t> keys{13, 37, 42, 88};
  CHECK(f.insert(keys.data(), keys.data() + keys.size()));
  auto probe = f.probe();
  std::vector<uint32_t> match_pos(keys.size(), 0);
  uint64_t c = probe(keys.data(), keys.data() + keys.size(), match_pos.data(), 0);
  CHECK(c == 4);
  CHECK(match_pos[0] == 0 && match_pos[1] == 1 && match_pos[2] == 2 && match_pos[3] == 3);
  std::vector<uint32_t> probe_keys{0, 1, 42, 2};
  c = probe(probe_keys.data(), probe_keys.data() + probe_keys.size(), match_pos.data(), 1000);
  CHECK(c == 1);
  CHECK(match_pos[0] == 1002);
  std::printf("PASS basic_build_probe_test\n");
}

static void class_api_test() {
  dtl::blocked_bloomfilter<uint64_t> bf(1ull << 20, 8, 8, 8);
  CHECK(bf.get_addressing_mode() == dtl::block_addressing::POWER_OF_TWO);
  CHECK(bf.size_in_bytes() == (1u << 17));
  uint64_t* words = nullptr;
  CHECK(hipMalloc(&words, bf.size_in_bytes()) == hipSuccess);
  CHECK(hipMemset(words, 0, bf.size_in_bytes()) == hipSuccess);
  std::vector<uint32_t> keys;
  for (uint32_t i = 0; i < 5000; ++i) keys.push_back(i * 2654435761u);
  uint32_t* dk = dev_copy(keys);
  bf.batch_insert(words, dk, static_cast<uint32_t>(keys.size()));
  uint32_t* dp = nullptr;
  CHECK(hipMalloc(&dp, keys.size() * 4) == hipSuccess);
  const uint64_t c = bf.batch_contains(words, dk, static_cast<uint32_t>(keys.size()), dp, 7);
  CHECK(c == keys.size());
  std::vector<uint32_t> pos(c);
  CHECK(hipMemcpy(pos.data(), dp, c * 4, hipMemcpyDeviceToHost) == hipSuccess);
  for (uint64_t i = 0; i < c; ++i) CHECK(pos[i] == i + 7);
  CHECK(bf.contains(words, keys[17]));
  bf.insert(words, 0xDEADBEEFu);
  CHECK(bf.contains(words, 0xDEADBEEFu));
  (void)hipFree(words);
  (void)hipFree(dk);
  (void)hipFree(dp);
  std::printf("PASS class_api_test\n");
}

static void invalid_config_test() {
  bool thrown = false;
  try {
    dtl::blocked_bloomfilter<uint32_t> bad(1024, 3, 1, 2);  // k % s != 0
  } catch (const std::invalid_argument&) {
    thrown = true;
  }
  CHECK(thrown);
  thrown = false;
  try {
    dtl::zoned_blocked_bloomfilter<uint32_t> bad(1024, 8, 4, 4);  // z >= w
  } catch (const std::invalid_argument&) {
    thrown = true;
  }
  CHECK(thrown);
  thrown = false;
  try {
    dtl::blocked_bloomfilter_config c;
    c.word_size = 2;
    dtl::filter::filter f(c, 1024);
  } catch (const std::runtime_error&) {
    thrown = true;
  }
  CHECK(thrown);
  std::printf("PASS invalid_config_test\n");
}

static void zoned_and_type_erased_test() {
  dtl::zbbf_32 z(10 * 4096, 6, 16, 2);
  CHECK(z.instance().get_addressing_mode() == dtl::block_addressing::MAGIC);
  uint64_t* d = nullptr;
  CHECK(hipMalloc(&d, z.size() * 8) == hipSuccess);
  CHECK(hipMemset(d, 0, z.size() * 8) == hipSuccess);
  std::vector<uint32_t> keys;
  for (uint32_t i = 1; i <= 4096; ++i) keys.push_back(i * 40503u);
  uint32_t* dk = dev_copy(keys);
  dtl::filter::filter_base& fb = z;
  CHECK(fb.batch_insert(d, dk, 4096));
  uint32_t* dp = nullptr;
  CHECK(hipMalloc(&dp, 4096 * 4) == hipSuccess);
  CHECK(fb.batch_contains(d, dk, 4096, dp, 0) == 4096);
  (void)hipFree(d);
  (void)hipFree(dk);
  (void)hipFree(dp);
  std::printf("PASS zoned_and_type_erased_test\n");
}

static void cuckoo_test() {
  dtl::cuckoofilter::config cc;
  cc.bits_per_tag = 16;
  cc.tags_per_bucket = 4;
  dtl::filter::filter f(cc, 1u << 20);
  std::vector<uint32_t> keys;
  for (uint32_t i = 1; i <= 15000; ++i) keys.push_back(i * 2246822519u);
  CHECK(f.insert(keys.data(), keys.data() + keys.size()));
  auto probe = f.probe();
  std::vector<uint32_t> pos(keys.size());
  CHECK(probe(keys.data(), keys.data() + keys.size(), pos.data(), 0) == keys.size());
  dtl::cuckoofilter::cuckoofilter raw(1u << 20, 16, 4);
  std::vector<uint32_t> table(raw.buffer_size_in_bytes() / 4 + 1, 0);
  CHECK(raw.batch_insert(table.data(), keys.data(), static_cast<uint32_t>(keys.size())));
  CHECK(raw.count_occupied_slots(table.data()) == keys.size());
  // GPU build of the same keys into a device table: every key found by the GPU probe
  void* d_table = nullptr;
  uint32_t* d_keys = nullptr;
  uint32_t* d_pos = nullptr;
  CHECK(hipMalloc(&d_table, raw.buffer_size_in_bytes() + 16) == hipSuccess);
  CHECK(hipMalloc(reinterpret_cast<void**>(&d_keys), keys.size() * 4) == hipSuccess);
  CHECK(hipMalloc(reinterpret_cast<void**>(&d_pos), keys.size() * 4) == hipSuccess);
  CHECK(hipMemset(d_table, 0, raw.buffer_size_in_bytes() + 16) == hipSuccess);
  CHECK(hipMemcpy(d_keys, keys.data(), keys.size() * 4, hipMemcpyHostToDevice) == hipSuccess);
  CHECK(raw.batch_insert_device(static_cast<uint32_t*>(d_table), d_keys, keys.size()));
  CHECK(raw.batch_contains(static_cast<const uint32_t*>(d_table), d_keys, static_cast<uint32_t>(keys.size()),
                           d_pos, 0) == keys.size());
  std::vector<uint32_t> gtable(table.size(), 0);
  CHECK(hipMemcpy(gtable.data(), d_table, raw.buffer_size_in_bytes(), hipMemcpyDeviceToHost) == hipSuccess);
  CHECK(raw.count_occupied_slots(gtable.data()) == keys.size());
  (void)hipFree(d_table);
  (void)hipFree(d_keys);
  (void)hipFree(d_pos);
  bool thrown = false;
  try {
    dtl::cuckoofilter::cuckoofilter bad(1u << 20, 7, 4);
  } catch (const std::invalid_argument&) {
    thrown = true;
  }
  CHECK(thrown);
  std::printf("PASS cuckoo_test\n");
}

// filter(n, tw): configuration chosen by the learned cost model (small
// calibration grid to keep the test short), explicit model and default
// instance loaded from $DTL_HIP_MODEL.
static void learned_config_test() {
  dtl_calib_params cp = {13, 22, 18, 1, 20, 0};
  auto cm = dtl::filter::model::cost_model::calibrate(&cp, true, 200000, 8);
  const dtl_bbf_choice ch = cm.choose(20000, 100);
  CHECK(ch.m >= 4 * 20000ull && ch.m <= 20 * 20000ull);
  dtl::filter::filter f(20000, 100, cm);
  std::vector<uint32_t> keys;
  for (uint32_t i = 1; i <= 20000; ++i) keys.push_back(i * 2654435761u);
  CHECK(f.insert(keys.data(), keys.data() + keys.size()));
  std::vector<uint32_t> pos(keys.size());
  auto probe = f.probe();
  CHECK(probe(keys.data(), keys.data() + keys.size(), pos.data(), 0) == keys.size());
  const char* path = "/tmp/dtl_hip_model_test.bin";
  cm.save(path);
  setenv("DTL_HIP_MODEL", path, 1);
  dtl::filter::filter g(20000, 100);  // default instance, loaded from the file
  CHECK(g.instance().size_in_bytes() == f.instance().size_in_bytes());
  CHECK(g.insert(keys.data(), keys.data() + 5000));
  auto probe_g = g.probe();
  CHECK(probe_g(keys.data(), keys.data() + 5000, pos.data(), 3) == 5000 && pos[0] == 3);
  std::remove(path);
  std::printf("PASS learned_config_test\n");
}

int main() {
  basic_build_probe_test();
  class_api_test();
  invalid_config_test();
  zoned_and_type_erased_test();
  cuckoo_test();
  learned_config_test();
  std::printf("ALL PASS\n");
  return 0;
}
