// Boxes over tagged nodes (host code): the clustering a regrid runs between
// tagging a level and defining the next finer level's boxes.
#pragma once

#include <cstdint>
#include <vector>

namespace uclbs {

struct ClusterBox {
  int lo[3], hi[3];  // inclusive, in the tags' index space
};

// tags: dense [z][y][x] bytes over lo..hi (inclusive), nonzero = tagged.
// Every tag is grown by `buffer` nodes (a cube, clipped to lo..hi), the grown
// tags coarsened by `blocking`, clustered by Berger and Rigoutsos' signature
// method (IEEE Trans. SMC 21(5), 1991) until each box holds at least
// `efficiency` of its nodes tagged, refined back, clipped to lo..hi and cut
// into pieces of at most `max_size` nodes per axis (0: no limit).  The boxes
// are disjoint and cover every grown tag; their faces lie on multiples of
// `blocking` except where clipped to lo..hi.  Returns nullptr, or what is wrong
// with the arguments.
const char* cluster_tags(const uint8_t* tags, const int32_t lo[3], const int32_t hi[3],
                         double efficiency, int blocking, int buffer, int max_size,
                         std::vector<ClusterBox>& out);

}  // namespace uclbs
