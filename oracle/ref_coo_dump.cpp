// Driver around the reference's own COO reader and virtual-graph builder
// (peel/shared/graph.cpp:17-99, peel/shared/virtual_graph.cpp:9-183),
// compiled from the sources where they lie.  TEST INFRASTRUCTURE ONLY.
//
// usage: ref_coo_dump <edge-file> <out.bin>
// out.bin (little-endian u32 words):
//   num_nodes, num_edges, numParts(lo), numParts(hi), hasZeroID,
//   nodePointer[num_nodes], edgeList[2*num_edges+num_nodes],
//   partNodePointer[numParts] as (node, part) pairs
#include "graph.h"
#include "virtual_graph.h"

#include <cstdio>

int main(int argc, char **argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s <edge-file> <out.bin>\n", argv[0]);
        return 2;
    }
    Graph graph(argv[1], false);
    graph.ReadGraph();
    bool zero = graph.hasZeroID;
    VirtualGraph vg(graph, 64);
    vg.MakeUGraph();
    FILE *f = std::fopen(argv[2], "wb");
    if (!f)
        return 3;
    unsigned long long np = (unsigned long long)vg.numParts;
    unsigned hdr[5] = {graph.num_nodes, graph.num_edges, (unsigned)(np & 0xffffffffu),
                       (unsigned)(np >> 32), zero ? 1u : 0u};
    std::fwrite(hdr, 4, 5, f);
    std::fwrite(vg.nodePointer, 4, graph.num_nodes, f);
    std::fwrite(vg.edgeList, 4, 2ull * graph.num_edges + graph.num_nodes, f);
    std::fwrite(vg.partNodePointer, 8, np, f);
    std::fclose(f);
    return 0;
}
