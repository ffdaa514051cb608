#include "xtcore_api.h"

#include <cstring>

namespace xtamd {
namespace render {

void tile_t::write(size_t x, size_t y, const float rgba[4]) {
    const size_t i = 4 * ((y - m_y0) * width() + (x - m_x0));
    std::memcpy(&m_data[i], rgba, 4 * sizeof(float));
}

void tile_t::read(size_t x, size_t y, float rgba[4]) const {
    const size_t i = 4 * ((y - m_y0) * width() + (x - m_x0));
    std::memcpy(rgba, &m_data[i], 4 * sizeof(float));
}

void tile_t::init() {
    m_data.assign(4 * width() * height(), 0.f);
    if (m_on_init) m_on_init->handle_event(this);
}

void tile_t::submit() {
    if (m_on_done) m_on_done->handle_event(this);
}

// Row-major tiles, clipped at the image edge (tile.cc:67-94).
void segment_framebuffer(Tileset &tiles, size_t width, size_t height, size_t tile_size) {
    tiles.clear();
    if (tile_size == 0) return;
    const size_t tx = width / tile_size + (width % tile_size ? 1 : 0);
    const size_t ty = height / tile_size + (height % tile_size ? 1 : 0);
    for (size_t j = 0; j < ty; ++j)
        for (size_t i = 0; i < tx; ++i) {
            size_t x0 = i * tile_size, y0 = j * tile_size;
            size_t x1 = x0 + tile_size, y1 = y0 + tile_size;
            if (x1 > width) x1 = width;
            if (y1 > height) y1 = height;
            tiles.emplace_back(x0, y0, x1, y1);
        }
}

context_t::~context_t() { xt_scene_free(scene); }

int context_t::load(const std::string &path, const std::vector<std::string> &mods) {
    xt_scene_free(scene);
    scene = nullptr;
    std::vector<const char *> m;
    for (auto &s : mods) m.push_back(s.c_str());
    int rc = xt_scene_load(path.c_str(), m.data(), (int)m.size(), &scene);
    if (rc == XT_OK) rc = xt_scene_build(scene, nullptr);
    error = scene ? xt_scene_error(scene) : "out of memory";
    if (rc == XT_OK) return 0;
    xt_scene_free(scene);
    scene = nullptr;
    return rc == XT_E_IO && error.find("parse") != std::string::npos ? 2 : 1;
}

void context_t::init() { segment_framebuffer(tiles, params.width, params.height, params.tile_size); }

void assemble(std::vector<float> &rgba, const context_t &ctx) {
    rgba.assign(4 * ctx.params.width * ctx.params.height, 0.f);
    float px[4];
    for (const tile_t &t : ctx.tiles)
        for (size_t y = t.y0(); y < t.y1(); ++y)
            for (size_t x = t.x0(); x < t.x1(); ++x) {
                t.read(x, y, px);
                std::memcpy(&rgba[4 * (y * ctx.params.width + x)], px, sizeof px);
            }
}

}  // namespace render

namespace gpu {

int Integrator::setup_auxiliary() {
    if (!ctx || !ctx->scene) return m_status = XT_E_STATE;
    if (!m_dev) {
        m_status = xt_dev_create(&m_dev, m_device, 0);
        if (m_status) return m_status;
    }
    if (!m_uploaded) {
        m_status = xt_scene_upload(m_dev, ctx->scene, &m_scene_id);
        if (m_status) return m_status;
        m_uploaded = true;
    }
    return m_status = XT_OK;
}

void Integrator::clean_auxiliary() {
    if (m_dev) {
        if (m_uploaded) xt_scene_release(m_dev, m_scene_id);
        xt_dev_destroy(m_dev);
    }
    m_dev = nullptr;
    m_uploaded = false;
}

const char *Integrator::error() const { return m_dev ? xt_last_error(m_dev) : "no device"; }

// IIntegrator::render (integrator.cc:29-53): returns silently without a
// context or camera, like the reference; otherwise one device job covers
// every tile.
void Integrator::render() {
    if (!ctx || !ctx->scene) return;
    if (!m_uploaded && setup_auxiliary() != XT_OK) return;
    const render::params_t &p = ctx->params;
    xt_job job{};
    if (xt_scene_camera(ctx->scene, p.camera.c_str(), &job.camera) != XT_OK) {
        m_status = XT_E_NOTFOUND;
        return;
    }
    job.scene_id = m_scene_id;
    job.integrator = (uint32_t)m_kind;
    job.width = (uint32_t)p.width;
    job.height = (uint32_t)p.height;
    job.samples = (uint32_t)p.samples;
    job.aa = (uint32_t)p.aa;
    job.aa_mode = p.sample_distribution == render::SAMPLE_DISTRIBUTION_RANDOM ? XT_AA_RANDOM : XT_AA_GRID;
    job.rdepth = (uint32_t)p.rdepth;
    job.accel = (uint32_t)p.accel;
    job.seed = p.seed;
    std::vector<float> fb(4 * p.width * p.height);
    for (render::tile_t &t : ctx->tiles) t.init();
    m_status = xt_render(m_dev, &job, fb.data(), &m_stats);
    if (m_status != XT_OK) return;
    for (render::tile_t &t : ctx->tiles) {
        for (size_t y = t.y0(); y < t.y1(); ++y)
            for (size_t x = t.x0(); x < t.x1(); ++x) t.write(x, y, &fb[4 * (y * p.width + x)]);
        t.submit();
    }
}

}  // namespace gpu
}  // namespace xtamd
