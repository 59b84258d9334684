// QuadDecomposition's maximum matching (QuadDecomposition.cpp:28-309), shared
// by the host decomposition (rz_quad.cpp) and the device one (rz_quad.hip,
// which builds the candidate graph on the GPU).  Host only.
#pragma once

#include <stdint.h>

namespace rzq {

// The candidate graph of n triangles as flat adjacency lists (vertex v's
// neighbours are adj[start[v] .. start[v + 1]), in the reference's link
// order); writes each triangle's partner, or -1, to mate.  The reference's
// greedy pass and augmenting-path search, step for step.  May throw
// std::bad_alloc.
void match(uint32_t n, const uint32_t* start, const int* adj, int* mate);

}  // namespace rzq
