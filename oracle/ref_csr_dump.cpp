// Driver around the reference's deduplicating CSR reader
// (peel/src/graph.cpp:58-133, used by peelone_dyn), compiled from the
// source where it lies.  TEST INFRASTRUCTURE ONLY.  Run it from an empty
// working directory: the reader probes ./output/serialized-<name>
// (src/graph.cpp:9-54) and must not find a stale cache there.
//
// usage: ref_csr_dump <edge-file> <out.bin>
// out.bin (u32 words): V, E, degrees[V], neighbors_offset[V+1], neighbors[E]
#include "graph.h"

#include <cstdio>

int main(int argc, char **argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s <edge-file> <out.bin>\n", argv[0]);
        return 2;
    }
    Graph g{std::string(argv[1])};
    FILE *f = std::fopen(argv[2], "wb");
    if (!f)
        return 3;
    unsigned hdr[2] = {g.V, g.E};
    std::fwrite(hdr, 4, 2, f);
    std::fwrite(g.degrees, 4, g.V, f);
    std::fwrite(g.neighbors_offset, 4, g.V + 1ull, f);
    std::fwrite(g.neighbors, 4, g.E, f);
    std::fclose(f);
    return 0;
}
