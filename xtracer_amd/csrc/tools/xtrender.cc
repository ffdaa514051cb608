// Headless driver over the xtcore-shaped C++ API: loads a .scn, renders with
// the MI355X integrator and writes the assembled framebuffer. Flags follow
// the reference CLI (src/frontend/cli/argdefs.h:7-17) where they apply:
//   xtrender scene.scn [-renderer stencil|depth|ao|pathtracer|normal|emission|wireframe] [-res WxH]
//            [-samples N] [-rdepth N] [-aa N] [-cam NAME] [-mod a.b.c:v]
//            [-accel kdtree|octree] [-seed N] [-outdir DIR] [-out FILE]
// With neither -out nor -outdir the image goes to <scene>_<renderer>.png in
// the current directory, with -outdir to DIR/<scene>_<renderer>.png, as the
// reference CLI names it (src/frontend/cli/xtracer.cc:125-146). -out picks
// the format by extension (.png .bmp .tga .hdr .exr .pfm .ppm, or .raw for
// the RGBA floats as they are).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../host/xtcore_api.h"

using namespace xtamd;

static int usage() {
    std::fprintf(stderr,
                 "usage: xtrender scene.scn [-renderer stencil|depth|ao|pathtracer|normal|uv|emission|wireframe] [-res WxH]\n"
                 "       [-samples N] [-rdepth N] [-aa N] [-tile_size N] [-threads N] [-cam NAME] [-mod a.b.c:v]\n"
                 "       [-accel kdtree|octree] [-seed N]\n"
                 "       [-outdir DIR] [-out FILE(.png|.bmp|.tga|.hdr|.exr|.pfm|.ppm|.raw)]\n");
    return 1;
}

int main(int argc, char **argv) {
    if (argc < 2) return usage();
    std::string scene = argv[1], out, outdir, renderer = "pathtracer";
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
        else if (a == "-tile_size") ctx.params.tile_size = std::strtoul(v, nullptr, 10);
        else if (a == "-threads") ctx.params.threads = std::strtoul(v, nullptr, 10);  // CPU threads: no effect here
        else if (a == "-cam") ctx.params.camera = v;
        else if (a == "-mod") mods.push_back(v);
        else if (a == "-accel") ctx.params.accel = std::strcmp(v, "octree") ? XT_ACCEL_KDTREE : XT_ACCEL_OCTREE;
        else if (a == "-seed") ctx.params.seed = (uint32_t)std::strtoul(v, nullptr, 10);
        else if (a == "-out") out = v;
        else if (a == "-outdir") outdir = v;
        else return usage();
        ++i;
    }
    gpu::Kind kind = renderer == "stencil" ? gpu::STENCIL
                     : renderer == "depth" ? gpu::DEPTH
                     : renderer == "ao" ? gpu::AO
                     : renderer == "normal" ? gpu::NORMAL
                     : renderer == "uv" ? gpu::UV
                     : renderer == "emission" ? gpu::EMISSION
                     : renderer == "wireframe" ? gpu::WIREFRAME
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
    const size_t W = ctx.params.width, H = ctx.params.height;
    if (out.empty()) {
        std::string base = scene.substr(scene.find_last_of('/') == std::string::npos ? 0 : scene.find_last_of('/') + 1);
        base = base.substr(0, base.find_last_of('.'));
        if (!outdir.empty() && outdir.back() != '/') outdir += '/';
        out = outdir + base + "_" + renderer + ".png";
    }
    if (out.size() > 4 && out.substr(out.size() - 4) == ".raw") {
        FILE *f = std::fopen(out.c_str(), "wb");
        const bool ok = f && std::fwrite(rgba.data(), sizeof(float), rgba.size(), f) == rgba.size();
        if (f) std::fclose(f);
        if (!ok) {
            std::fprintf(stderr, "failed to write %s\n", out.c_str());
            return 5;
        }
    } else if (int wr = xt_image_write(out.c_str(), rgba.data(), (uint32_t)W, (uint32_t)H)) {
        std::fprintf(stderr, "failed to export %s (%d)\n", out.c_str(), wr);
        return 5;
    }
    const xt_job_stats &st = in.stats();
    std::printf("%s %zux%zu: %llu queries in %.3f ms (%.1f Mrays/s)\n", renderer.c_str(), W, H,
                (unsigned long long)st.queries, st.ms_total, st.ms_total > 0 ? st.queries / st.ms_total / 1e3 : 0.0);
    return 0;
}
