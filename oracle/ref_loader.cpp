// ref_loader.cpp — drives the reference harness's own graph Builder
// (utilities/builder.hpp, reader.hpp) on one .mtx / .sg file and writes the
// CSR it builds as raw little-endian uint64: nodes, edges, offsets[nodes+1],
// destinations[edges].  TEST INFRASTRUCTURE ONLY: compiled from the
// reference's own headers into oracle/_ref/ by oracle/Makefile, used by
// tests/golden/make_loader_golden.py to pin the library's loaders.
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <string>

#include "builder.hpp"

static int dump(const CSRGraph<SpB_Index, SpB_Index> &g, const char *path)
{
    uint64_t n = (uint64_t)g.num_nodes(), m = (uint64_t)g.num_edges();
    std::ofstream out(path, std::ios::binary);
    out.write(reinterpret_cast<const char *>(&n), 8);
    out.write(reinterpret_cast<const char *>(&m), 8);
    out.write(reinterpret_cast<const char *>(g.out_offsets()), (std::streamsize)(8 * (n + 1)));
    out.write(reinterpret_cast<const char *>(g.out_destValues()), (std::streamsize)(8 * m));
    out.close();
    return out ? 0 : 1;
}

int main(int argc, char **argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: ref_loader <graph.mtx|graph.sg> <out.bin>\n");
        return 2;
    }
    // The harness's one-argument Builder takes .sg / .wsg only
    // (builder.hpp:63-79); MatrixMarket input goes through the unweighted
    // two-argument form (builder.hpp:81-100).  The graph points into its
    // builder's offset array, so the builder stays alive while the graph is
    // written, as in the harness (test_spmv_vnec.cpp:142-143).
    std::string path = argv[1];
    bool mtx = path.size() >= 4 && path.compare(path.size() - 4, 4, ".mtx") == 0;
    int rc;
    if (mtx) {
        Builder<SpB_Index, SpB_Index> b(path, false);
        CSRGraph<SpB_Index, SpB_Index> g = b.MakeGraph();
        rc = dump(g, argv[2]);
    } else {
        Builder<SpB_Index, SpB_Index> b(path);
        CSRGraph<SpB_Index, SpB_Index> g = b.MakeGraph();
        rc = dump(g, argv[2]);
    }
    std::fflush(stdout);
    std::_Exit(rc); // the output is written; skip the reference objects' teardown
}
