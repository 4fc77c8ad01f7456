// Driver (ours) around the reference's own readers, compiled against
// code/graph.hpp and code/update.hpp where they lie.  Prints what
// graph::loadGraph (code/graph.hpp:202-223) and graph::parseUpdates
// (code/graph.hpp:185-200 -> code/update.hpp:52-89) read, so the engine's
// readers can be checked against them.  Test infrastructure only.
//   load_dump graph <file.bin>    -> "n m" then row, col, cap, rev lines
//   load_dump updates <file.bin>  -> "size" then one "type src dst w" per record
#include "graph.hpp"
#include <cstdio>
#include <cstring>

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  if (std::strcmp(argv[1], "graph") == 0) {
    graph g(argv[2]);
    g.loadGraph();
    const int n = g.num_nodes(), m = g.num_edges();
    std::printf("%d %d\n", n, m);
    for (int i = 0; i <= n; ++i) std::printf("%d%c", g.indexofNodes[i], i == n ? '\n' : ' ');
    for (int i = 0; i < m; ++i) std::printf("%d%c", g.edgeList[i], i + 1 == m ? '\n' : ' ');
    for (int i = 0; i < m; ++i) std::printf("%d%c", g.getEdgeLen()[i], i + 1 == m ? '\n' : ' ');
    for (int i = 0; i < m; ++i) std::printf("%d%c", g.parallel_edge[i], i + 1 == m ? '\n' : ' ');
    return 0;
  }
  if (std::strcmp(argv[1], "updates") == 0) {
    graph g(argv[2]);
    std::vector<update> u = g.parseUpdates(argv[2]);
    std::printf("%zu\n", u.size());
    for (const update& x : u) std::printf("%d %d %d %d\n", (int)x.type, x.source, x.destination, x.weight);
    return 0;
  }
  return 2;
}
