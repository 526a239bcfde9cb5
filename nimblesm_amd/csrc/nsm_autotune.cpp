// nsm_autotune.cpp -- process and file cache of the element-tile choice
// (see nsm_autotune.h).
#include "nsm_autotune.h"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>

namespace nsm {
namespace {

std::mutex                 g_mutex;
std::map<std::string, int> g_cache;
bool                       g_file_loaded = false;

const char* cache_file() { return std::getenv("NSM_AUTOTUNE_CACHE"); }

void load_file_locked()
{
  if (g_file_loaded) return;
  g_file_loaded = true;
  const char* path = cache_file();
  if (!path || !*path) return;
  std::ifstream in(path);
  std::string   line;
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    std::string        key;
    int                v = 0;
    if (ss >> key >> v)
      for (int c : kTileCandidates)
        if (c == v) g_cache[key] = v;
  }
}

void append_file(const std::string& key, int v)
{
  const char* path = cache_file();
  if (!path || !*path) return;
  std::ofstream out(path, std::ios::app);
  out << key << ' ' << v << '\n';
}

}  // namespace

int autotune_cached(const std::string& key)
{
  std::lock_guard<std::mutex> lk(g_mutex);
  load_file_locked();
  auto it = g_cache.find(key);
  return it == g_cache.end() ? 0 : it->second;
}

int autotune_tiles(const std::string& key, const std::function<double(int)>& time_of,
                   std::vector<std::pair<int, double>>* timings)
{
  if (timings) timings->clear();
  const int cached = autotune_cached(key);
  if (cached) return cached;
  int    best   = 0;
  double best_t = 0.0;
  for (int c : kTileCandidates) {
    const double t = time_of(c);
    if (timings) timings->emplace_back(c, t);
    if (t < 0.0) continue;
    if (!best || t < best_t) {
      best   = c;
      best_t = t;
    }
  }
  if (!best) return 0;
  std::lock_guard<std::mutex> lk(g_mutex);
  g_cache[key] = best;
  append_file(key, best);
  return best;
}

}  // namespace nsm
