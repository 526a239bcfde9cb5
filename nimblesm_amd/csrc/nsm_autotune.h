// nsm_autotune.h -- element-tile size chosen from timed device steps
// (internal).
//
// The tiled mixed-precision kernel runs one thread per tile element, so the
// tile size is also the workgroup size (64, 128, 256 or 512 threads: one to
// eight waves).  Large meshes fill the device at any size and favour 512
// (fewer shared nodes, shorter boundary pass); small meshes and ensembles of
// small meshes leave most CUs idle with 512-element tiles.  The choice for a
// problem shape is made once per process by timing a few steps with every
// candidate, kept in memory, and also kept in a text file when
// NSM_AUTOTUNE_CACHE names one (one "key tile_elems" line per shape).
#pragma once

#include <functional>
#include <string>
#include <utility>
#include <vector>

namespace nsm {

constexpr int kTileCandidates[] = {64, 128, 256, 512};

// The tile size for `key`: cached, else the candidate with the smallest
// time_of(candidate) (seconds per step; a negative value marks a candidate
// that cannot be used).  `timings` receives what was measured (empty when the
// choice came from a cache).  Returns 0 when no candidate can be used.
int autotune_tiles(const std::string& key, const std::function<double(int)>& time_of,
                   std::vector<std::pair<int, double>>* timings);

// The cached choice for key (0: none).
int autotune_cached(const std::string& key);

}  // namespace nsm
