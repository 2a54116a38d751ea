// BFS on the advance engine, written against the graph_one API the way the reference's apps are
// (a ComponentX subclass for the host-side status, an AdvanceFunctor for the per-edge rule, Run).
// The reference ships no BFS app; this one uses the same CLI and output format as apps/sssp_adv.cu:
//   bfs_adv -i graph.adj --src S [-o out.txt] [-d] ...   output lines "vid-depth" (depth -1: unreached)
#include <chrono>
#include <cstdio>
#include <limits>

#include "GraphOne/graph_one.h"

using namespace graph_one;

constexpr arch_t arch = arch_t::hip;
constexpr uint32_t kUnreached = std::numeric_limits<uint32_t>::max();

struct bfs_hstatus_t {
    vid_t src;
    DenseVec<arch, uint32_t>& depth;
};

struct bfs_dstatus_t {
    uint32_t level;
    uint32_t* depth;
};

template <typename graph_t>
struct BFSComp : ComponentX<graph_t, bfs_hstatus_t, bfs_dstatus_t> {
    using comp_t = ComponentX<graph_t, bfs_hstatus_t, bfs_dstatus_t>;
    using comp_t::comp_t;

    void Init() override {
        archi::fill<arch>(this->h_status.depth.begin(), this->h_status.depth.end(), kUnreached);
        this->h_status.depth.set(this->h_status.src, 0);
    }
    void BeforeEngine() override { ++this->d_status.level; }
};

struct BFSFunctor : AdvanceFunctor<vid_t, eid_t, float, bfs_dstatus_t> {
    // The first visitor of dst claims it; the claim is the filter, so no filter() is needed.
    __ONE_DEV_INL__ static bool advance(const vid_t&, const vid_t& dst, const eid_t&, const float&, bfs_dstatus_t& d) {
        return atomicCAS(&d.depth[dst], kUnreached, d.level) == kUnreached;
    }
};

int main(int argc, char* argv[]) {
    CLI::App app;
    std::string input_graph, output_path;
    bool reorder_vid = false;
    LoadEdgeOpts opts;
    add_common_args(app, input_graph, output_path, reorder_vid, opts);
    vid_t src = 0;
    app.add_option("--src", src, "source vertex of the BFS")->required();
    CLI11_PARSE(app, argc, argv);

    GraphLoader<vstart_t::FROM_0_TO_0> loader(reorder_vid);
    auto cache = loader.LoadEdgesFromTxt<float>(input_graph, opts);
    auto g = graph::build<arch, AdvanceViews>(cache);
    using graph_t = decltype(g);
    if (!loader.ReorderedVid(src) || src >= g.num_vertices()) LOG_ERROR("src vertex \"", src, "\" does not exist");

    DenseVec<arch, uint32_t> depth(g.num_vertices());
    bfs_hstatus_t h{src, depth};
    bfs_dstatus_t d{0, depth.data()};
    SpDblFrontier<arch, vid_t> frontier(g.num_vertices(), {src});
    BFSComp<graph_t> comp(g, h, d);

    const auto t0 = std::chrono::high_resolution_clock::now();
    Run<BFSFunctor>(comp, frontier);
    const auto t1 = std::chrono::high_resolution_clock::now();
    const auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count();
    printx("Elapsed time: ", ms, " ms");
    if (!output_path.empty()) {
        FILE* fp = fopen(output_path.c_str(), "w");
        if (!fp) LOG_ERROR("open output file failed");
        fprintf(fp, "Elapsed time: %lld ms\n", static_cast<long long>(ms));
        auto hd = depth.to<arch_t::cpu>();
        for (vid_t i = 0; i < hd.size(); ++i) fprintf(fp, "%u-%d\n", i, hd[i] == kUnreached ? -1 : static_cast<int>(hd[i]));
        fclose(fp);
    }
    return 0;
}
