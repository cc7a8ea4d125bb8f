// kc_cli.cpp -- drop-in host programs for the reference executables.
//
// One binary, dispatched on argv[0] (symlinks gpp, peelone, peelone_dyn,
// HistoCore, CntCore, NbrCore).  Each keeps the reference's argument
// contract and stdout lines and calls the device only through libkcore's C
// ABI (include/kcore.h):
//   gpp / peelone / HistoCore / CntCore / NbrCore : argument.cpp:36-66
//     (exactly one argument), shared/graph.cpp:17-99 reader, output lines of
//     gppeel.cu:231-241 and util.cpp:6-21 (Min K starts at 10).
//   peelone_dyn : main.cu:65-98 (dedup reader, STD / Maxdegree, ten
//     repetitions of peel_dyn and peelone_dyn, drop min and max).
// KC_SAVE_RESULTS=<path> additionally writes "i core" lines in the format of
// utility::SaveResults (util.cpp:70-89), which the reference never reaches.
// KC_CSR_CACHE=<dir> reads the edge file through the content-keyed binary
// CSR cache in <dir> (the reference's ./output/serialized-<name> text cache,
// peel/src/graph.cpp:9-54); stdout is the same either way.
#include "kcore.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

static std::string base_name(const char *p)
{
    const char *s = std::strrchr(p, '/');
    return s ? s + 1 : p;
}

static std::string help_string()
{
    // argument::GenerateHelpString (argument.cpp:68-82) for canHaveSource =
    // canHaveIters = false
    std::string str = "\nRequired arguments:";
    str += "\n    [--input]: Input graph file. E.g., --input FacebookGraph.txt";
    str += "\nOptional arguments";
    str += "\n    [--output]: Output file for results. E.g., --output results.txt";
    str += "\n    [--device]: Select GPU device (default: 0). E.g., --device 1";
    str += "\n\n";
    return str;
}

static void print_max_min(const std::vector<uint32_t> &core)
{
    uint32_t mx = 0, mn = 10; // util.cpp:8-9
    for (uint32_t c : core) {
        mx = std::max(mx, c);
        mn = std::min(mn, c);
    }
    std::printf(" Max K = %u\n Min K = %u\n", mx, mn);
}

static kc_status read_graph(const char *path, int reader, kc_host_graph **out)
{
    const char *dir = std::getenv("KC_CSR_CACHE");
    if (dir && *dir)
        return kc_hg_read_cached(path, reader, dir, out, nullptr);
    return reader == KC_READER_COO ? kc_hg_read_coo(path, out) : kc_hg_read_dedup(path, out);
}

static int die(const char *what)
{
    std::fprintf(stderr, "%s: %s\n", what, kc_last_error());
    return 1;
}

static void maybe_save(const std::vector<uint32_t> &core)
{
    const char *p = std::getenv("KC_SAVE_RESULTS");
    if (!p || !*p)
        return;
    std::printf("Saving the results into the following file:\n>> %s\n", p);
    if (kc_write_cores(p, core.data(), (uint32_t)core.size()) == KC_OK)
        std::printf("Done saving.\n");
    else
        std::fprintf(stderr, "saving failed: %s\n", kc_last_error());
}

static int run_vgraph_program(const std::string &prog, const char *path)
{
    kc_method m = prog == "gpp"         ? KC_GPP
                  : prog == "peelone"   ? KC_PEELONE
                  : prog == "HistoCore" ? KC_HISTO
                  : prog == "CntCore"   ? KC_CNT
                                        : KC_NBR;
    std::printf("Reading the input graph from the following file:\n>> %s\n", path);
    kc_host_graph *hg = nullptr;
    if (read_graph(path, KC_READER_COO, &hg) != KC_OK)
        return die("read");
    kc_graph_info info;
    kc_hg_info(hg, &info);
    std::printf("Done reading.\n");
    std::printf("Number of nodes = %u\n", info.V);
    std::printf("Number of edges = %llu\n", (unsigned long long)info.edges_read);
    kc_ctx *ctx = nullptr;
    if (kc_ctx_create(0, &ctx) != KC_OK)
        return die("device");
    kc_graph *g = nullptr;
    if (kc_graph_upload_hg(ctx, hg, 4, &g) != KC_OK)
        return die("upload");
    kc_opts o;
    std::memset(&o, 0, sizeof o);
    o.gpp_compat = 1; // gppeel reports isolated vertices as 1 (gppeel.cu:182-191)
    o.part_size = 64; // argument.h:24
    std::vector<uint32_t> core(info.V);
    kc_stats st;
    if (kc_decompose(g, m, &o, core.data(), &st) != KC_OK)
        return die("decompose");
    std::printf("Number of iterations = %llu\n", (unsigned long long)st.iterations);
    std::printf("Processing finished in %g (ms).\n", st.device_ms);
    print_max_min(core);
    maybe_save(core);
    kc_graph_free(g);
    kc_ctx_destroy(ctx);
    kc_hg_free(hg);
    return 0;
}

// main.cu:28-49: ten runs, drop the largest and smallest, average eight
static void rep_simulation(int (*kern)(kc_ref_graph *, uint32_t *), kc_ref_graph &g, bool &ok)
{
    float sum = 0;
    const int rep = 10;
    unsigned max_time = 0, min_time = 0xffffffffu;
    for (int i = 0; i < rep; i++) {
        int t = kern(&g, nullptr);
        if (t < 0) {
            ok = false;
            return;
        }
        std::printf("%ums ", (unsigned)t);
        max_time = std::max(max_time, (unsigned)t);
        min_time = std::min(min_time, (unsigned)t);
        sum += (float)t;
    }
    sum = sum - max_time - min_time;
    std::printf(" ave: %g\n", sum * 1.0 / (rep - 2));
}

static int run_peelone_dyn(int argc, char **argv)
{
    if (argc < 2) {
        std::printf("./kcore file deviceID Repeattime\n");
        return 255; // exit(-1)
    }
    std::printf("Graph loading Started... \n");
    kc_host_graph *hg = nullptr;
    if (read_graph(argv[1], KC_READER_DEDUP, &hg) != KC_OK)
        return die("read");
    kc_graph_info info;
    kc_hg_info(hg, &info);
    std::vector<uint32_t> off(info.V + 1ull), deg(info.V);
    for (uint64_t v = 0; v <= info.V; v++)
        off[v] = (uint32_t)info.rowptr[v];
    for (uint32_t v = 0; v < info.V; v++)
        deg[v] = off[v + 1] - off[v];
    kc_ref_graph g;
    g.V = info.V;
    g.E = (uint32_t)info.nnz;
    g.neighbors = const_cast<uint32_t *>(info.col);
    g.neighbors_offset = off.data();
    g.degrees = deg.data();
    g.kmax = 0;
    g.dmax = info.dmax;
    std::printf("V: %u\n", g.V);
    std::printf("E: %u\n", g.E / 2);
    // STDdegrees (main.cu:51-63): population standard deviation
    double sum = 0;
    for (uint32_t d : deg)
        sum += d;
    double mean = g.V ? sum / g.V : 0.0, sq = 0;
    for (uint32_t d : deg)
        sq += (d - mean) * (d - mean);
    std::printf("\nSTD: \n%g\n", g.V ? std::sqrt(sq / g.V) : 0.0);
    uint32_t mx = 0;
    for (uint32_t d : deg)
        mx = std::max(mx, d);
    std::printf(" Maxdegree = %u\n", mx);
    bool ok = true;
    std::printf("ICDE23_peel: ");
    rep_simulation(kc_peel_dyn, g, ok);
    if (!ok)
        return die("peel_dyn");
    std::printf("Kmax: %u\n", g.kmax);
    std::printf("Our_Paper_peel: ");
    rep_simulation(kc_peelone_dyn, g, ok);
    if (!ok)
        return die("peelone_dyn");
    std::printf("Kmax: %u\n", g.kmax);
    if (std::getenv("KC_SAVE_RESULTS")) {
        std::vector<uint32_t> core(g.V);
        if (kc_peelone_dyn(&g, core.data()) < 0)
            return die("peelone_dyn");
        maybe_save(core);
    }
    kc_hg_free(hg);
    return 0;
}

int main(int argc, char **argv)
{
    std::string prog = base_name(argv[0]);
    if (prog == "peelone_dyn")
        return run_peelone_dyn(argc, argv);
    if (prog != "gpp" && prog != "peelone" && prog != "HistoCore" && prog != "CntCore" &&
        prog != "NbrCore") {
        std::fprintf(stderr, "run as one of: gpp peelone peelone_dyn HistoCore CntCore NbrCore\n");
        return 2;
    }
    if (argc != 2) { // argument.cpp:40-43
        std::printf("%s", help_string().c_str());
        return 0;
    }
    return run_vgraph_program(prog, argv[1]);
}
