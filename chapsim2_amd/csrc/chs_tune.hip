#include "chs_tune.h"

#include <cstdio>
#include <map>
#include <mutex>
#include <vector>

namespace chs {

namespace {
struct Entry {
  int choice = -1;
  std::vector<float> us;
};
std::mutex g_mu;
std::map<std::string, Entry> g_cache;
std::string g_force_prefix;
int g_force_variant = -1;
}  // namespace

void tune_force(const std::string& prefix, int variant) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_force_prefix = prefix;
  g_force_variant = variant;
}

void tune_set(const std::string& key, int choice) {
  std::lock_guard<std::mutex> lk(g_mu);
  Entry e;
  e.choice = choice;
  g_cache[key] = e;
}

int tuned_choice(const std::string& key) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_cache.find(key);
  return it == g_cache.end() ? -1 : it->second.choice;
}

int forced_or_tuned(const std::string& key, int nvar) {
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_force_variant >= 0 && key.compare(0, g_force_prefix.size(), g_force_prefix) == 0)
      return g_force_variant < nvar ? g_force_variant : -1;
  }
  return tuned_choice(key);
}

int tune(const std::string& key, int nvar, hipStream_t s, const std::function<bool(int)>& launch) {
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_force_variant >= 0 && key.compare(0, g_force_prefix.size(), g_force_prefix) == 0)
      return g_force_variant < nvar ? g_force_variant : -1;
  }
  const int have = tuned_choice(key);
  if (have >= 0) return have;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) return -1;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess) return -1;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return -1;
  }
  Entry ent;
  float best = 0.0f;
  for (int v = 0; v < nvar; ++v) {
    float tv = -1.0f;
    if (launch(v)) {  // warm-up
      for (int r = 0; r < 3; ++r) {
        float ms = 0.0f;
        if (hipEventRecord(e0, s) != hipSuccess || !launch(v) || hipEventRecord(e1, s) != hipSuccess ||
            hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
          tv = -1.0f;
          break;
        }
        if (tv < 0.0f || ms < tv) tv = ms;
      }
    }
    ent.us.push_back(tv < 0.0f ? -1.0f : 1000.0f * tv);
    if (tv >= 0.0f && (ent.choice < 0 || tv < best)) {
      ent.choice = v;
      best = tv;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipGetLastError();
  if (ent.choice < 0) return -1;
  std::lock_guard<std::mutex> lk(g_mu);
  g_cache[key] = ent;
  return ent.choice;
}

std::string tune_report() {
  std::lock_guard<std::mutex> lk(g_mu);
  std::string out;
  char buf[64];
  for (const auto& kv : g_cache) {
    out += kv.first + "=" + std::to_string(kv.second.choice) + " (";
    for (size_t i = 0; i < kv.second.us.size(); ++i) {
      std::snprintf(buf, sizeof(buf), "%s%.1f", i ? "/" : "", kv.second.us[i]);
      out += buf;
    }
    out += " us)\n";
  }
  return out;
}

}  // namespace chs
