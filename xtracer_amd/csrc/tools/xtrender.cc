// Headless driver over the xtcore-shaped C++ API: loads a .scn, renders with
// the MI355X integrator and writes the assembled framebuffer as a PFM (RGB,
// float) or raw RGBA floats. Flags follow the reference CLI
// (src/frontend/cli/argdefs.h:7-17) where they apply:
//   xtrender scene.scn [-renderer stencil|depth|ao|pathtracer] [-res WxH]
//            [-samples N] [-rdepth N] [-aa N] [-cam NAME] [-mod a.b.c:v]
//            [-accel kdtree|octree] [-seed N] [-out FILE(.pfm|.raw)]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../host/xtcore_api.h"

using namespace xtamd;

static int usage() {
    std::fprintf(stderr,
                 "usage: xtrender scene.scn [-renderer stencil|depth|ao|pathtracer] [-res WxH] [-samples N]\n"
                 "       [-rdepth N] [-aa N] [-cam NAME] [-mod a.b.c:v] [-accel kdtree|octree] [-seed N] [-out FILE]\n");
    return 1;
}

int main(int argc, char **argv) {
    if (argc < 2) return usage();
    std::string scene = argv[1], out = "out.pfm", renderer = "pathtracer";
    std::vector<std::string> mods;
    render::context_t ctx;
    for (int i = 2; i < argc; ++i) {
        std::string a = argv[i];
        const char *v = i + 1 < argc ? argv[i + 1] : nullptr;
        if (!v) return usage();
        if (a == "-renderer") renderer = v;
        else if (a == "-res") {
            unsigned w, h;
            if (std::sscanf(v, "%ux%u", &w, &h) != 2) return usage();
            ctx.params.width = w;
            ctx.params.height = h;
        } else if (a == "-samples") ctx.params.samples = std::strtoul(v, nullptr, 10);
        else if (a == "-rdepth") ctx.params.rdepth = std::strtoul(v, nullptr, 10);
        else if (a == "-aa") ctx.params.aa = std::strtoul(v, nullptr, 10);
        else if (a == "-cam") ctx.params.camera = v;
        else if (a == "-mod") mods.push_back(v);
        else if (a == "-accel") ctx.params.accel = std::strcmp(v, "octree") ? XT_ACCEL_KDTREE : XT_ACCEL_OCTREE;
        else if (a == "-seed") ctx.params.seed = (uint32_t)std::strtoul(v, nullptr, 10);
        else if (a == "-out") out = v;
        else return usage();
        ++i;
    }
    gpu::Kind kind = renderer == "stencil" ? gpu::STENCIL
                     : renderer == "depth" ? gpu::DEPTH
                     : renderer == "ao" ? gpu::AO
                     : gpu::PATHTRACER;
    int rc = ctx.load(scene, mods);
    if (rc) {
        std::fprintf(stderr, "failed to load %s (%d): %s\n", scene.c_str(), rc, ctx.error.c_str());
        return 2;
    }
    ctx.init();
    gpu::Integrator in(kind);
    in.setup(ctx);
    if (in.setup_auxiliary() != XT_OK) {
        std::fprintf(stderr, "device setup failed: %s\n", in.error());
        return 3;
    }
    in.render();
    if (in.status() != XT_OK) {
        std::fprintf(stderr, "render failed: %s\n", in.error());
        return 4;
    }
    std::vector<float> rgba;
    render::assemble(rgba, ctx);
    FILE *f = std::fopen(out.c_str(), "wb");
    if (!f) return 5;
    const size_t W = ctx.params.width, H = ctx.params.height;
    if (out.size() > 4 && out.substr(out.size() - 4) == ".raw") {
        std::fwrite(rgba.data(), sizeof(float), rgba.size(), f);
    } else {
        std::fprintf(f, "PF\n%zu %zu\n-1.0\n", W, H);
        std::vector<float> row(3 * W);
        for (size_t y = H; y-- > 0;) {  // PFM stores rows bottom to top
            for (size_t x = 0; x < W; ++x)
                for (int c = 0; c < 3; ++c) row[3 * x + c] = rgba[4 * (y * W + x) + c];
            std::fwrite(row.data(), sizeof(float), row.size(), f);
        }
    }
    std::fclose(f);
    const xt_job_stats &st = in.stats();
    std::printf("%s %zux%zu: %llu queries in %.3f ms (%.1f Mrays/s)\n", renderer.c_str(), W, H,
                (unsigned long long)st.queries, st.ms_total, st.ms_total > 0 ? st.queries / st.ms_total / 1e3 : 0.0);
    return 0;
}
