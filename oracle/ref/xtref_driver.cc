// TEST INFRASTRUCTURE ONLY. A C-callable driver around the REAL reference
// xtcore, compiled from /root/reference by oracle/ref/Makefile into
// oracle/_ref/libxtref.so. Tests use it to pin the oracle and the loader;
// nothing in the product links it. Modelled on the GUI's render driver
// (src/frontend/gui/workspace.cc:55-81), since the reference CLI cannot
// render (its integrator selection is commented out).
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <string>

#include <omp.h>

#include <xtcore/context.h>
#include <xtcore/integrator.h>
#include <xtcore/math/plane.h>
#include <xtcore/math/sphere.h>
#include <xtcore/math/triangle.h>
#include <xtcore/mesh.h>
#include <xtcore/octree.h>
#include <xtcore/parseutil.h>
#include <xtcore/strpool.h>
#include <xtcore/log.h>
#include <xtcore/util/raygraph.h>
#include <xtcore/aa.h>
#include <xtcore/camera/perspective.h>
#include <xtcore/emitter.h>
#include <xtcore/material.h>
#include <nmath/prng.h>
#include <xtcore/integrator/wireframe/integrator.h>
#include <nimg/img.h>
#include <nimg/pixmap.h>

using namespace xtcore;

namespace {
bool g_pool = false;
// Per-thread query counters (one cache line each), bumped by the link-time
// wrapper of Scene::intersection below.
struct alignas(64) Counter {
    uint64_t n;
};
Counter g_queries[512];
}  // namespace

// The Makefile links with --wrap on the mangled Scene::intersection symbol,
// so every call the integrators make lands here first; the reference's own
// code is unchanged.
#define XR_ISECT _ZN6xtcore5Scene12intersectionERKNS_3RayERNS_12hit_record_tE
#define XR_CAT2(a, b) a##b
#define XR_CAT(a, b) XR_CAT2(a, b)
extern "C" bool XR_CAT(__real_, XR_ISECT)(void *self, const void *ray, void *hit);
extern "C" bool XR_CAT(__wrap_, XR_ISECT)(void *self, const void *ray, void *hit) {
    g_queries[omp_get_thread_num() & 511].n++;
    return XR_CAT(__real_, XR_ISECT)(self, ray, hit);
}

extern "C" {

void *xr_load(const char *path, int *status) {
    if (!g_pool) {
        xtcore::pool::str::init();
        xtcore::Log::handle().echo(false);
        g_pool = true;
    }
    render::context_t *ctx = new render::context_t;
    int rc = io::scn::load(&ctx->scene, path, nullptr);
    if (status) *status = rc;
    if (rc) {
        delete ctx;
        return nullptr;
    }
    return ctx;
}

void xr_free(void *h) { delete (render::context_t *)h; }

uint32_t xr_object_count(void *h) { return (uint32_t)((render::context_t *)h)->scene.m_objects.size(); }

// Object i in scan order: name, geometry kind (0 plane, 1 sphere, 2
// triangle, 3 mesh, -1 missing) and its parameters.
int xr_object(void *h, uint32_t i, char *name, int name_len, int *kind, double *param) {
    Scene &s = ((render::context_t *)h)->scene;
    if (i >= s.m_objects.size()) return 1;
    auto it = s.m_objects.begin();
    std::advance(it, i);
    std::strncpy(name, pool::str::get(it->first), name_len - 1);
    name[name_len - 1] = 0;
    auto g = s.m_surface.find(it->second->surface);
    *kind = -1;
    if (g == s.m_surface.end() || !g->second) return 0;
    asset::ISurface *surf = g->second;
    if (auto *p = dynamic_cast<surface::Plane *>(surf)) {
        *kind = 0;
        param[0] = p->normal.x;
        param[1] = p->normal.y;
        param[2] = p->normal.z;
        param[3] = p->offset;
    } else if (auto *sp = dynamic_cast<surface::Sphere *>(surf)) {
        *kind = 1;
        param[0] = sp->origin.x;
        param[1] = sp->origin.y;
        param[2] = sp->origin.z;
        param[3] = sp->radius;
    } else if (auto *t = dynamic_cast<surface::Triangle *>(surf)) {
        *kind = 2;
        for (int k = 0; k < 3; ++k) {
            param[3 * k] = t->v[k].x;
            param[3 * k + 1] = t->v[k].y;
            param[3 * k + 2] = t->v[k].z;
        }
    } else if (dynamic_cast<surface::Mesh *>(surf)) {
        *kind = 3;
    }
    return 0;
}

// Scene::intersection for each ray (origin and direction assigned as given,
// bypassing the normalising Ray constructor). obj: scan index or -1.
void xr_trace(void *h, uint64_t n, const double *rays, double *t, int32_t *obj, double *point, double *normal) {
    Scene &s = ((render::context_t *)h)->scene;
    for (uint64_t i = 0; i < n; ++i) {
        Ray r;
        r.origin = nmath::Vector3f(rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]);
        r.direction = nmath::Vector3f(rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]);
        hit_record_t hr;
        bool hit = s.intersection(r, hr);
        t[i] = hit ? hr.t : INFINITY;
        obj[i] = -1;
        if (hit) {
            int k = 0;
            for (auto it = s.m_objects.begin(); it != s.m_objects.end(); ++it, ++k)
                if (it->first == hr.id_object) obj[i] = k;
        }
        if (point) {
            point[3 * i] = hr.point.x;
            point[3 * i + 1] = hr.point.y;
            point[3 * i + 2] = hr.point.z;
        }
        if (normal) {
            normal[3 * i] = hr.normal.x;
            normal[3 * i + 1] = hr.normal.y;
            normal[3 * i + 2] = hr.normal.z;
        }
    }
}

// Builds Octree<Triangle> exactly as Mesh::build_octree does (mesh.cc:70-113)
// from n_tris triangles (9 doubles each) and returns, per ray, the index of
// the triangle Octree::intersection picks (-1 on a miss) and its t. The index
// rides in the triangle's unused texcoord slot.
void xr_octree_trace(uint64_t n_tris, const double *tris, uint64_t n, const double *rays, int64_t *index, double *t,
                     uint64_t *node_count) {
    Octree<surface::Triangle> oct;
    for (uint64_t i = 0; i < n_tris; ++i) {
        surface::Triangle p;
        for (int k = 0; k < 3; ++k)
            p.v[k] = nmath::Vector3f(tris[9 * i + 3 * k], tris[9 * i + 3 * k + 1], tris[9 * i + 3 * k + 2]);
        p.tc[0] = nmath::Vector2f((double)i, 0.0);
        p.calc_aabb();
        oct.add(p.aabb, p);
    }
    oct.max_items_per_node(LIMITS_MAX_ITEMS_PER_NODE);
    oct.max_depth(LIMITS_MAX_DEPTH);
    oct.build();
    (void)node_count;
    for (uint64_t i = 0; i < n; ++i) {
        Ray r;
        r.origin = nmath::Vector3f(rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]);
        r.direction = nmath::Vector3f(rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]);
        hit_record_t hr;
        OctreeItem<surface::Triangle> *it = oct.intersection(r, &hr);
        index[i] = it ? (int64_t)it->data.tc[0].x : -1;
        t[i] = it ? hr.t : INFINITY;
    }
}

// Renders with a reference integrator (0 stencil, 1 depth, 2 ao, 3 pathtracer,
// 4 normal, 5 uv, 6 emission, 7 wireframe) and assembles the framebuffer into rgba (W*H*4 floats).
// aa_mode: 0 GRID, 1 RANDOM (params_t::sample_distribution).
int xr_render(void *h, int integrator, int width, int height, int samples, int aa, int aa_mode, int rdepth,
              int threads, unsigned seed, float *rgba) {
    render::context_t *ctx = (render::context_t *)h;
    render::IIntegrator *in = nullptr;
    switch (integrator) {
        case 0: in = new integrator::stencil::Integrator; break;
        case 1: in = new integrator::depth::Integrator; break;
        case 2: in = new integrator::ao::Integrator; break;
        case 3: in = new integrator::pathtracer::Integrator; break;
        case 4: in = new integrator::normal::Integrator; break;
        case 5: in = new integrator::uv::Integrator; break;
        case 6: in = new integrator::emission::Integrator; break;
        case 7: in = new integrator::wireframe::Integrator; break;
        default: return 1;
    }
    render::params_t p;
    p.width = width;
    p.height = height;
    p.samples = samples;
    p.aa = aa;
    p.rdepth = rdepth;
    p.threads = threads;
    p.sample_distribution = aa_mode == 1 ? antialiasing::SAMPLE_DISTRIBUTION_RANDOM
                                         : antialiasing::SAMPLE_DISTRIBUTION_GRID;
    p.camera = ctx->scene.m_cameras.empty() ? 0 : ctx->scene.m_cameras.begin()->first;
    ctx->params = p;
    ctx->init();
    in->setup(*ctx);
    srand(seed);
    in->render();
    nimg::Pixmap fb;
    render::assemble(fb, *ctx);
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            const nimg::ColorRGBAf &c = fb.pixel_ro(x, y);
            float *o = rgba + 4 * ((size_t)y * width + x);
            o[0] = c.r();
            o[1] = c.g();
            o[2] = c.b();
            o[3] = c.a();
        }
    delete in;
    return 0;
}

// Renders only every `stride`-th tile (row-major order) and returns the
// wall time of render() and the number of Scene::intersection calls it made.
int xr_render_tiles(void *h, int integrator, int width, int height, int samples, int aa, int rdepth, int threads,
                    unsigned seed, int stride, double *seconds, uint64_t *queries) {
    render::context_t *ctx = (render::context_t *)h;
    render::IIntegrator *in = nullptr;
    switch (integrator) {
        case 0: in = new integrator::stencil::Integrator; break;
        case 1: in = new integrator::depth::Integrator; break;
        case 2: in = new integrator::ao::Integrator; break;
        case 3: in = new integrator::pathtracer::Integrator; break;
        case 4: in = new integrator::normal::Integrator; break;
        case 5: in = new integrator::uv::Integrator; break;
        case 6: in = new integrator::emission::Integrator; break;
        case 7: in = new integrator::wireframe::Integrator; break;
        default: return 1;
    }
    render::params_t p;
    p.width = width;
    p.height = height;
    p.samples = samples;
    p.aa = aa;
    p.rdepth = rdepth;
    p.threads = threads;
    p.camera = ctx->scene.m_cameras.empty() ? 0 : ctx->scene.m_cameras.begin()->first;
    ctx->params = p;
    ctx->init();
    render::Tileset keep;
    for (size_t i = 0; i < ctx->tiles.size(); i += (size_t)(stride > 0 ? stride : 1)) keep.push_back(ctx->tiles[i]);
    ctx->tiles.swap(keep);
    in->setup(*ctx);
    srand(seed);
    for (auto &c : g_queries) c.n = 0;
    auto t0 = std::chrono::steady_clock::now();
    in->render();
    auto t1 = std::chrono::steady_clock::now();
    uint64_t q = 0;
    for (auto &c : g_queries) q += c.n;
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    if (queries) *queries = q;
    delete in;
    return 0;
}

// The ray graph of the last xr_render through the reference's own
// render::assemble(raygraph_t&) and raygraph::write (context.cc:54-61,
// util/raygraph.cc:8-78), as the GUI exports it (action.cc:65-85).
int xr_write_raygraph(void *h, const char *path) {
    render::context_t *ctx = (render::context_t *)h;
    raygraph::raygraph_t rg;
    render::assemble(rg, *ctx);
    return raygraph::write(path, rg);
}

// Primary rays of the scene's first camera (the one xr_render uses) at n
// sample coordinates xy (float pairs), through ICamera::get_primary_ray.
int xr_camera_rays(void *h, uint64_t n, const float *xy, float width, float height, double *o, double *d) {
    render::context_t *ctx = (render::context_t *)h;
    if (ctx->scene.m_cameras.empty()) return 1;
    asset::ICamera *cam = ctx->scene.get_camera(ctx->scene.m_cameras.begin()->first);
    if (!cam) return 1;
    for (uint64_t i = 0; i < n; ++i) {
        Ray r = cam->get_primary_ray(xy[2 * i], xy[2 * i + 1], width, height);
        o[3 * i] = r.origin.x;
        o[3 * i + 1] = r.origin.y;
        o[3 * i + 2] = r.origin.z;
        d[3 * i] = r.direction.x;
        d[3 * i + 1] = r.direction.y;
        d[3 * i + 2] = r.direction.z;
    }
    return 0;
}

// Thin-lens primary rays of the scene's first camera with the lens draws
// recorded. glibc rand() is one process-global sequence, so after
// srand(seed + i) the two prng_c(-a/2, a/2) draws get_primary_ray makes
// (perspective.cc:77-78) are reproduced exactly by seeding again and calling
// the same prng_c here. Every translation unit holds its own copy of
// prng_c's "initialized" flag; both copies are touched first so that neither
// re-seeds from time(NULL) later. jitter receives 2 doubles per ray.
int xr_thin_lens_rays(void *h, uint64_t n, const float *xy, float width, float height, unsigned seed, double *o,
                      double *d, double *jitter) {
    render::context_t *ctx = (render::context_t *)h;
    if (ctx->scene.m_cameras.empty()) return 1;
    asset::ICamera *cam = ctx->scene.get_camera(ctx->scene.m_cameras.begin()->first);
    camera::Perspective *p = dynamic_cast<camera::Perspective *>(cam);
    if (!p) return 1;
    (void)p->get_primary_ray(0.f, 0.f, width, height);
    (void)nmath::prng_c(0.0, 1.0);
    const nmath::scalar_t half = p->aperture / 2.f;  // perspective.cc:76
    for (uint64_t i = 0; i < n; ++i) {
        srand(seed + (unsigned)i);
        Ray r = cam->get_primary_ray(xy[2 * i], xy[2 * i + 1], width, height);
        srand(seed + (unsigned)i);
        jitter[2 * i] = nmath::prng_c(-half, half);
        jitter[2 * i + 1] = nmath::prng_c(-half, half);
        o[3 * i] = r.origin.x;
        o[3 * i + 1] = r.origin.y;
        o[3 * i + 2] = r.origin.z;
        d[3 * i] = r.direction.x;
        d[3 * i + 1] = r.direction.y;
        d[3 * i + 2] = r.direction.z;
    }
    return 0;
}

// antialiasing::gen_samples_random (aa.cc:34-48) for n pixels at `level`,
// with its draws recorded the same way: per pixel, srand(seed + i), the
// reference's sample set, then srand(seed + i) again and the same number of
// prng_c(0, 1) draws. coords and draws receive 2 * level doubles per pixel,
// weights level doubles per pixel, in the order the set pops them.
int xr_aa_random(uint64_t n, const float *pixels, unsigned level, unsigned seed, double *coords, double *weights,
                 double *draws) {
    {
        antialiasing::sample_set_t warm;
        antialiasing::gen_samples_random(warm, nmath::Vector2f(0, 0), 1);
    }
    (void)nmath::prng_c(0.0, 1.0);
    for (uint64_t i = 0; i < n; ++i) {
        antialiasing::sample_set_t set;
        srand(seed + (unsigned)i);
        antialiasing::gen_samples_random(set, nmath::Vector2f(pixels[2 * i], pixels[2 * i + 1]), level);
        if (set.count() != level) return 1;
        for (unsigned k = 0; k < level; ++k) {
            antialiasing::sample_rgba_t s;
            set.pop(s);
            coords[2 * (i * level + k)] = s.coords.x;
            coords[2 * (i * level + k) + 1] = s.coords.y;
            weights[i * level + k] = s.weight;
        }
        srand(seed + (unsigned)i);
        for (unsigned k = 0; k < 2 * level; ++k) draws[2 * i * level + k] = nmath::prng_c(0.0, 1.0);
    }
    return 0;
}

// Writes rgba (W*H*4 floats) through the reference's own image writers
// (lib/nimg/img.cc): 0 png, 1 bmp, 2 tga, 3 hdr, 4 exr.
// IMaterial::shade of object i's material for n queries: q holds 16 doubles
// per query (hit point, normal, texcoord, camera position, emitter position,
// object index), ei the emitter intensity and io the intensity argument (3
// floats each, io updated in place); ret gets shade()'s return value, or -1
// for an object without a material.
int xr_shade(void *h, uint64_t n, const double *q, const float *ei, float *io, int *ret) {
    Scene &s = ((render::context_t *)h)->scene;
    camera::Perspective cam;
    for (uint64_t k = 0; k < n; ++k) {
        const double *a = q + 16 * k;
        const uint32_t oi = (uint32_t)a[15];
        ret[k] = -1;
        if (oi >= s.m_objects.size()) continue;
        auto it = s.m_objects.begin();
        std::advance(it, oi);
        const asset::IMaterial *m = s.get_material(it->first);
        if (!m) continue;
        hit_record_t hr;
        hr.point = nmath::Vector3f(a[0], a[1], a[2]);
        hr.normal = nmath::Vector3f(a[3], a[4], a[5]);
        hr.texcoord = nmath::Vector3f(a[6], a[7], a[8]);
        cam.position = nmath::Vector3f(a[9], a[10], a[11]);
        asset::emitter_t em;
        em.position = nmath::Vector3f(a[12], a[13], a[14]);
        em.intensity = nimg::ColorRGBf(ei[3 * k], ei[3 * k + 1], ei[3 * k + 2]);
        nimg::ColorRGBf in(io[3 * k], io[3 * k + 1], io[3 * k + 2]);
        ret[k] = m->shade(in, &cam, &em, hr) ? 1 : 0;
        io[3 * k] = in.r();
        io[3 * k + 1] = in.g();
        io[3 * k + 2] = in.b();
    }
    return 0;
}

int xr_save_image(int fmt, const char *path, const float *rgba, int w, int h) {
    nimg::Pixmap pm;
    if (pm.init((size_t)w, (size_t)h)) return 1;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const float *p = rgba + 4 * ((size_t)y * w + x);
            pm.pixel((size_t)x, (size_t)y) = nimg::ColorRGBAf(p[0], p[1], p[2], p[3]);
        }
    switch (fmt) {
        case 0: return nimg::io::save::png(path, pm);
        case 1: return nimg::io::save::bmp(path, pm);
        case 2: return nimg::io::save::tga(path, pm);
        case 3: return nimg::io::save::hdr(path, pm);
        case 4: return nimg::io::save::exr(path, pm);
        default: return 1;
    }
}

// Loads `path` through the reference's own loader (nimg::io::load::image,
// lib/nimg/img.cc:17-66, stb_image underneath). With rgba NULL only the size
// is returned; otherwise rgba receives w*h*4 floats, row 0 at the top.
int xr_load_image(const char *path, float *rgba, int *w, int *h) {
    nimg::Pixmap pm;
    const int rc = nimg::io::load::image(path, pm);
    if (rc) return rc;
    *w = (int)pm.width();
    *h = (int)pm.height();
    if (!rgba) return 0;
    for (size_t y = 0; y < pm.height(); ++y)
        for (size_t x = 0; x < pm.width(); ++x) {
            const nimg::ColorRGBAf &c = pm.pixel_ro(x, y);
            float *p = rgba + 4 * (y * pm.width() + x);
            p[0] = c.r();
            p[1] = c.g();
            p[2] = c.b();
            p[3] = c.a();
        }
    return 0;
}

}  // extern "C"
