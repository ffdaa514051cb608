// C++ renderer interface mirroring xtcore's (src/xtcore/integrator.h:10-24,
// context.h:18-51, tile.h:23-59), built on the C ABI. A frontend written
// against xtcore's params_t / tile_t / context_t / IIntegrator keeps its shape:
//
//   xtamd::render::context_t ctx;
//   ctx.load("scene.scn");                     // io::scn::load
//   ctx.params.width = 1920; ...               // params_t, same defaults
//   ctx.init();                                // segment_framebuffer
//   xtamd::gpu::Integrator in(xtamd::gpu::PATHTRACER);
//   in.setup(ctx); in.setup_auxiliary();       // device + resident scene
//   in.render();                               // all tiles, one device job
//   xtamd::render::assemble(rgba, ctx);        // W*H RGBA float
//
// Differences that follow from the device path: render() is virtual (the GPU
// integrator renders every tile in one launch sequence and then fires each
// tile's on_init / on_done handlers in tile order, from the calling thread),
// and params_t::threads is ignored.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/xtracer_amd.h"

namespace xtamd {
namespace render {

enum TILE_ORDER { TILE_ORDER_SCANLINE, TILE_ORDER_RANDOM, TILE_ORDER_RADIAL_IN, TILE_ORDER_RADIAL_OUT };
enum SAMPLE_DISTRIBUTION { SAMPLE_DISTRIBUTION_GRID, SAMPLE_DISTRIBUTION_RANDOM };

// Defaults of src/xtcore/context.cc:3-27.
struct params_t {
    size_t width = 500, height = 500, threads = 0, samples = 1, aa = 1, rdepth = 100, tile_size = 32;
    TILE_ORDER tile_order = TILE_ORDER_RANDOM;
    SAMPLE_DISTRIBUTION sample_distribution = SAMPLE_DISTRIBUTION_GRID;
    std::string camera;  // name; empty: the first camera in name order
    uint32_t seed = 1;
    int accel = XT_ACCEL_KDTREE;
};

struct tile_t;
struct tile_event_handler_t {
    virtual ~tile_event_handler_t() = default;
    virtual void handle_event(tile_t *tile) = 0;
};

// A rectangle [x0, x1) x [y0, y1) of the framebuffer with its own FP32 RGBA
// pixels (tile.cc:8-94).
struct tile_t {
    tile_t(size_t x0, size_t y0, size_t x1, size_t y1) : m_x0(x0), m_y0(y0), m_x1(x1), m_y1(y1) {}
    size_t x0() const { return m_x0; }
    size_t x1() const { return m_x1; }
    size_t y0() const { return m_y0; }
    size_t y1() const { return m_y1; }
    size_t width() const { return m_x1 - m_x0; }
    size_t height() const { return m_y1 - m_y0; }
    void setup_handler_on_init(tile_event_handler_t *h) { m_on_init = h; }
    void setup_handler_on_done(tile_event_handler_t *h) { m_on_done = h; }
    void write(size_t x, size_t y, const float rgba[4]);
    void read(size_t x, size_t y, float rgba[4]) const;
    void init();    // zero the pixels, fire on_init
    void submit();  // fire on_done

private:
    size_t m_x0, m_y0, m_x1, m_y1;
    std::vector<float> m_data;
    tile_event_handler_t *m_on_init = nullptr;
    tile_event_handler_t *m_on_done = nullptr;
};
typedef std::vector<tile_t> Tileset;

void segment_framebuffer(Tileset &tiles, size_t width, size_t height, size_t tile_size);

struct context_t {
    context_t() = default;
    ~context_t();
    context_t(const context_t &) = delete;
    context_t &operator=(const context_t &) = delete;

    // io::scn::load: 0 ok, 1 failure, 2 parse error (parseutil.cc:697-714).
    int load(const std::string &path, const std::vector<std::string> &mods = {});
    void init();  // tiles from params (context.cc:29-33)

    xt_scene *scene = nullptr;
    Tileset tiles;
    params_t params;
    std::string error;
};

// Copies every tile into a W*H RGBA float framebuffer (context.cc:38-52).
void assemble(std::vector<float> &rgba, const context_t &ctx);

class IIntegrator {
public:
    virtual ~IIntegrator() = default;
    void setup(context_t &ctx) { this->ctx = &ctx; }
    virtual void render() = 0;
    virtual int setup_auxiliary() { return 0; }
    virtual void clean_auxiliary() {}
    context_t *ctx = nullptr;
};

}  // namespace render

namespace gpu {

enum Kind { STENCIL = XT_INTEGRATOR_STENCIL, DEPTH = XT_INTEGRATOR_DEPTH, AO = XT_INTEGRATOR_AO,
            PATHTRACER = XT_INTEGRATOR_PATHTRACER, NORMAL = XT_INTEGRATOR_NORMAL, UV = XT_INTEGRATOR_UV,
            EMISSION = XT_INTEGRATOR_EMISSION, WIREFRAME = XT_INTEGRATOR_WIREFRAME };

// The MI355X integrator: setup_auxiliary() creates the device and uploads the
// scene (the role the reference's Embree integrator gives it), render()
// renders every tile as one device job, clean_auxiliary() releases both.
class Integrator : public render::IIntegrator {
public:
    explicit Integrator(Kind kind, int hip_device = 0) : m_kind(kind), m_device(hip_device) {}
    ~Integrator() override { clean_auxiliary(); }
    int setup_auxiliary() override;
    void clean_auxiliary() override;
    void render() override;
    const xt_job_stats &stats() const { return m_stats; }
    int status() const { return m_status; }
    const char *error() const;

private:
    Kind m_kind;
    int m_device;
    xt_dev *m_dev = nullptr;
    uint32_t m_scene_id = 0;
    bool m_uploaded = false;
    xt_job_stats m_stats{};
    int m_status = XT_OK;
};

}  // namespace gpu
}  // namespace xtamd
