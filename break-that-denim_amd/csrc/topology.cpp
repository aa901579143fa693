// Detached-piece detection on the host (src/simulation/gpu/topology.ts:142-238): a breadth-first
// flood of the active span graph from the mounted nodes, then from each grip anchor not yet reached;
// cloth nodes it does not reach become detached. Detached nodes (from before, or freed now; not
// retired) are then labelled into pieces connected through active spans, each piece's nodes
// contiguous in the order the flood met them.
//
// Per node, `bits`: bit 0 / bit 1 = the active span of family u / v the node owns, bit 2 = not cloth
// (detached or retired), bit 3 = retired (state.wgsl:16-27, pack).
#include <cstdint>
#include <vector>

#include "../../include/bd.h"

extern "C" int bd_topology_search(int N, int layer_count, const int32_t* nx, const int32_t* ny, const int32_t* offset,
                                  const uint8_t* bits, const uint8_t* mounted, const int32_t* anchors, int anchor_count,
                                  uint8_t* freed, uint32_t* piece_of, uint32_t* piece_start, uint32_t* piece_nodes,
                                  int32_t* piece_count) {
  if (N < 0 || layer_count <= 0 || !nx || !ny || !offset || !bits || !mounted || !freed || !piece_of || !piece_start ||
      !piece_nodes || !piece_count || (anchor_count > 0 && !anchors))
    return BD_ERR_INVALID;
  // lattice row length, column and row of every node
  std::vector<int32_t> stride(N, 0), column(N, 0), row(N, 0), rows(N, 0);
  long total = 0;
  for (int l = 0; l < layer_count; l++) {
    if (nx[l] <= 0 || ny[l] <= 0 || offset[l] < 0 || (long)offset[l] + (long)nx[l] * ny[l] > N) return BD_ERR_INVALID;
    for (int k = 0; k < nx[l] * ny[l]; k++) {
      const int n = offset[l] + k;
      stride[n] = nx[l];
      column[n] = k % nx[l];
      row[n] = k / nx[l];
      rows[n] = ny[l];
    }
    total += (long)nx[l] * ny[l];
  }
  if (total != N) return BD_ERR_INVALID;
  for (int a = 0; a < anchor_count; a++)
    if (anchors[a] < 0 || anchors[a] >= N) return BD_ERR_INVALID;

  // neighbours of n through active spans (it owns the spans toward +u and +v; a span bit at the
  // lattice's edge names no neighbour and is ignored)
  auto neighbours = [&](int n, int out[4]) {
    int c = 0;
    const int s = stride[n];
    if ((bits[n] & 1) && column[n] + 1 < s) out[c++] = n + 1;
    if ((bits[n] & 2) && row[n] + 1 < rows[n]) out[c++] = n + s;
    if (column[n] > 0 && (bits[n - 1] & 1)) out[c++] = n - 1;
    if (row[n] > 0 && (bits[n - s] & 2)) out[c++] = n - s;
    return c;
  };

  // flood from the mounts, then from each anchor not yet reached (topology.ts:169-190)
  std::vector<uint8_t> visited(N, 0);
  std::vector<int32_t> queue;
  queue.reserve(N);
  auto push = [&](int n) {
    if (visited[n] || (bits[n] & 4)) return;
    visited[n] = 1;
    queue.push_back(n);
  };
  size_t head = 0;
  auto flood = [&]() {
    int nb[4];
    while (head < queue.size()) {
      const int n = queue[head++];
      const int c = neighbours(n, nb);
      for (int q = 0; q < c; q++) push(nb[q]);
    }
  };
  for (int n = 0; n < N; n++)
    if (mounted[n]) push(n);
  flood();
  for (int a = 0; a < anchor_count; a++) {
    const int n = anchors[a];
    if (visited[n] || (bits[n] & 4)) continue;
    push(n);
    flood();
  }
  int freed_count = 0;
  for (int n = 0; n < N; n++) {
    freed[n] = (!visited[n] && !(bits[n] & 4)) ? 1 : 0;
    freed_count += freed[n];
  }

  // pieces of the detached nodes through active spans (topology.ts:196-224)
  auto detached = [&](int n) { return ((bits[n] & 4) && !(bits[n] & 8)) || freed[n]; };
  for (int n = 0; n < N; n++) piece_of[n] = 0xFFFFFFFFu;
  int pieces = 0;
  uint32_t tail = 0;
  piece_start[0] = 0;
  int nb[4];
  for (int seed = 0; seed < N; seed++) {
    if (piece_of[seed] != 0xFFFFFFFFu || !detached(seed)) continue;
    const uint32_t p = (uint32_t)pieces;
    uint32_t at = tail;
    piece_of[seed] = p;
    piece_nodes[tail++] = (uint32_t)seed;
    while (at < tail) {
      const int n = (int)piece_nodes[at++];
      const int c = neighbours(n, nb);
      for (int q = 0; q < c; q++) {
        const int m = nb[q];
        if (piece_of[m] != 0xFFFFFFFFu || !detached(m)) continue;
        piece_of[m] = p;
        piece_nodes[tail++] = (uint32_t)m;
      }
    }
    pieces++;
    piece_start[pieces] = tail;
  }
  *piece_count = pieces;
  return freed_count;
}
