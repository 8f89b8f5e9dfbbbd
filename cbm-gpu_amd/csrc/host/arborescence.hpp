// Minimum spanning arborescence (optimum branching) by Tarjan's contraction
// method with meldable heaps: Tarjan, "Finding optimum branchings",
// Networks 7 (1977), with the correction of Camerini, Fratta and Maffioli,
// Networks 9 (1979).  O(m log n) time.
//
// Heaps are leftist trees with a lazy additive tag (the reduced-cost update
// of a contracted cycle costs O(1)); super-nodes are tracked by a union-find
// with union by size and an undo log, so the branching can be expanded back
// cycle by cycle at the end.
//
// Determinism: a heap orders edges by (reduced weight, edge id), so among
// edges of equal reduced weight the one created first wins; start vertices
// are visited in ascending id.  The same input therefore always yields the
// same parent array on every machine.
#pragma once

#include <cstdint>
#include <vector>

namespace cbmh {

struct ArborescenceResult {
    int64_t weight = 0;            // total weight of the chosen edges
    std::vector<int32_t> in_edge;  // chosen edge id entering each vertex, -1 for the root
    bool spanning = false;         // false if some vertex is unreachable from the root
    bool consistent = false;       // sum of picked reduced weights equals the final weight
};

// n vertices 0..n-1, root `root`; edge e goes src[e] -> dst[e] with weight w[e] >= 0.
ArborescenceResult min_arborescence(int32_t n, int32_t root, const int32_t *src, const int32_t *dst,
                                    const int32_t *w, int64_t n_edges);

}  // namespace cbmh
