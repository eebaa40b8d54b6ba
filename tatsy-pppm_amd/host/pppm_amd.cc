// pppm_amd.cc -- the reference-shaped C++ classes over the C ABI.
#include "pppm_amd.h"

#include <cmath>
#include <cstring>
#include <mutex>

namespace pppm_amd {

// ---- Vector3D (sources/vector3d.cc:44-199) ---------------------------------
double Vector3D::dot(const Vector3D& u, const Vector3D& v) { return u._x * v._x + u._y * v._y + u._z * v._z; }
Vector3D Vector3D::cross(const Vector3D& u, const Vector3D& v) {
    return Vector3D(u._y * v._z - u._z * v._y, u._z * v._x - u._x * v._z, u._x * v._y - u._y * v._x);
}
Vector3D Vector3D::minimum(const Vector3D& u, const Vector3D& v) {
    return Vector3D(std::fmin(u._x, v._x), std::fmin(u._y, v._y), std::fmin(u._z, v._z));
}
Vector3D Vector3D::maximum(const Vector3D& u, const Vector3D& v) {
    return Vector3D(std::fmax(u._x, v._x), std::fmax(u._y, v._y), std::fmax(u._z, v._z));
}
double Vector3D::norm() const { return std::sqrt(squaredNorm()); }
Vector3D Vector3D::normalized() const { return *this / (norm() + 1.0e-12); }  // vector3d.cc:63-65
Vector3D& Vector3D::operator+=(const Vector3D& v) { _x += v._x; _y += v._y; _z += v._z; return *this; }
Vector3D& Vector3D::operator-=(const Vector3D& v) { _x -= v._x; _y -= v._y; _z -= v._z; return *this; }
Vector3D& Vector3D::operator*=(double s) { _x *= s; _y *= s; _z *= s; return *this; }
Vector3D operator+(const Vector3D& u, const Vector3D& v) { Vector3D r = u; r += v; return r; }
Vector3D operator-(const Vector3D& u, const Vector3D& v) { Vector3D r = u; r -= v; return r; }
Vector3D operator*(const Vector3D& u, const Vector3D& v) { return Vector3D(u.x() * v.x(), u.y() * v.y(), u.z() * v.z()); }
Vector3D operator*(const Vector3D& v, double s) { Vector3D r = v; r *= s; return r; }
Vector3D operator*(double s, const Vector3D& v) { return v * s; }
Vector3D operator/(const Vector3D& v, double s) { return v * (1.0 / s); }

void BBox::merge(const Vector3D& v) {
    _posMin = Vector3D::minimum(_posMin, v);
    _posMax = Vector3D::maximum(_posMax, v);
}
void BBox::merge(const BBox& b) {
    merge(b._posMin);
    merge(b._posMax);
}
void BBox::merge(const Triangle& t) {
    for (int k = 0; k < 3; k++) merge(t.p(k));
}

// ---- BSDF ------------------------------------------------------------------
BSDF::BSDF() { std::memset(&_b, 0, sizeof(_b)); }
Vector3D BSDF::reflectance() const { return Vector3D(_b.reflectance[0], _b.reflectance[1], _b.reflectance[2]); }

static BSDF make_bsdf(int type, const Vector3D& r) {
    pppm_bsdf b;
    std::memset(&b, 0, sizeof(b));
    b.type = type;
    b.reflectance[0] = (float)r.x();
    b.reflectance[1] = (float)r.y();
    b.reflectance[2] = (float)r.z();
    b.ior = 1.5f;  // IOR_OBJECT, sources/common.h:16
    b.eta = 1.3f;
    return BSDF(b);
}
BSSRDF DipoleBSSRDF::factory(const Vector3D& sigma_a, const Vector3D& sigmap_s, double eta, double scale) {
    if (scale != 1.0) throw Error(PPPM_ERR_INVALID, "DipoleBSSRDF: only scale 1 is supported");
    return BSSRDF(sigma_a, sigmap_s, eta);
}

void BSDF::setBssrdf(const BSSRDF& bssrdf) {
    for (int k = 0; k < 3; k++) {
        _b.sigma_a[k] = (float)bssrdf.sigma_a()[k];
        _b.sigmap_s[k] = (float)bssrdf.sigmap_s()[k];
    }
    _b.eta = (float)bssrdf.eta();
    _b.type |= PPPM_BSDF_BSSRDF;
}

BSDF LambertianBRDF::factory(const Vector3D& r) { return make_bsdf(PPPM_BSDF_LAMBERTIAN, r); }
BSDF SpecularBRDF::factory(const Vector3D& r) { return make_bsdf(PPPM_BSDF_SPECULAR, r); }
BSDF RefractionBSDF::factory(const Vector3D& r) { return make_bsdf(PPPM_BSDF_REFRACTION, r); }

PerspectiveCamera::PerspectiveCamera(const Vector3D& center, const Vector3D& direction, const Vector3D& up,
                                     double fov, int w, int h, double sensitivity)
    : _sensitivity(sensitivity) {
    std::memset(&_c, 0, sizeof(_c));
    for (int d = 0; d < 3; d++) {
        _c.center[d] = center[d];
        _c.direction[d] = direction[d];
        _c.up[d] = up[d];
    }
    _c.fov_y_deg = fov;
    _c.width = w;
    _c.height = h;
}

Vector3D Image::operator()(int x, int y) const {
    size_t k = 3 * ((size_t)y * _w + x);
    return Vector3D(_rgb[k], _rgb[k + 1], _rgb[k + 2]);
}

// ---- Device ----------------------------------------------------------------
Device::Device(int device) {
    int rc = pppm_create(&_ctx, device);
    if (rc != PPPM_OK) throw Error(rc, pppm_last_error(nullptr));
}
Device::~Device() {
    if (_ctx) pppm_destroy(_ctx);
}
void Device::check(int rc) const {
    if (rc != PPPM_OK) throw Error(rc, pppm_last_error(_ctx));
}
std::shared_ptr<Device> Device::shared(int device) {
    static std::mutex m;
    static std::weak_ptr<Device> cache;
    std::lock_guard<std::mutex> g(m);
    std::shared_ptr<Device> d = cache.lock();
    if (!d) {
        d = std::make_shared<Device>(device);
        cache = d;
    }
    return d;
}

static std::vector<float> flatten(const std::vector<Triangle>& tris) {
    std::vector<float> v(9 * tris.size());
    for (size_t i = 0; i < tris.size(); i++)
        for (int k = 0; k < 3; k++)
            for (int d = 0; d < 3; d++) v[9 * i + 3 * k + d] = (float)tris[i].p(k)[d];
    return v;
}

static void fill_hits(const std::vector<pppm_hit>& hits, const std::vector<Ray>& rays,
                      const std::vector<Triangle>& tris, std::vector<Intersection>* out) {
    out->resize(rays.size());
    for (size_t i = 0; i < rays.size(); i++) {
        Intersection is;
        is.setObjectId(hits[i].tri);
        Hitpoint hp;
        if (hits[i].tri >= 0) {
            hp.setDistance(hits[i].t);
            hp.setPosition(rays[i].origin() + hits[i].t * rays[i].direction());
            hp.setNormal(tris[hits[i].tri].normal());
        }
        is.setHitpoint(hp);
        (*out)[i] = is;
    }
}

static void batch_intersect(void* ctx, const Device& dev, const std::vector<Ray>& rays,
                            std::vector<pppm_hit>* hits, float tmax = 1.0e30f) {
    std::vector<pppm_ray> r(rays.size());
    for (size_t i = 0; i < rays.size(); i++) {
        for (int d = 0; d < 3; d++) {
            r[i].origin[d] = (float)rays[i].origin()[d];
            r[i].dir[d] = (float)rays[i].direction()[d];
        }
        r[i].tmin = 0.0f;
        r[i].tmax = tmax;
    }
    hits->resize(rays.size());
    dev.check(pppm_intersect(ctx, r.data(), (int64_t)r.size(), hits->data()));
}

// ---- QBVHAccel -------------------------------------------------------------
QBVHAccel::QBVHAccel() : _dev(std::make_shared<Device>(0)) {}
QBVHAccel::QBVHAccel(std::shared_ptr<Device> dev) : _dev(std::move(dev)) {}

void QBVHAccel::construct(const std::vector<Triangle>& triangles) {
    _tris = triangles;
    std::vector<float> v = flatten(triangles);
    std::vector<int32_t> ids(triangles.size(), 0);
    pppm_bsdf b = LambertianBRDF::factory(Vector3D(0.5, 0.5, 0.5)).raw();
    pppm_light l;
    std::memset(&l, 0, sizeof(l));
    l.edge_u[0] = 1.0f;
    l.edge_v[2] = 1.0f;
    _dev->check(pppm_scene_upload(_dev->handle(), v.data(), ids.data(), (int64_t)triangles.size(), &b, 1, &l, 1));
    _dev->check(pppm_bvh_build(_dev->handle(), nullptr));
    _built = true;
}

int QBVHAccel::intersect(const Ray& ray, Hitpoint* hitpoint) const {
    if (!_built) throw Error(PPPM_ERR_STATE, "QBVHAccel::construct not called");
    std::vector<pppm_hit> h;
    batch_intersect(_dev->handle(), *_dev, std::vector<Ray>(1, ray), &h, (float)hitpoint->distance());
    if (h[0].tri < 0) return -1;
    hitpoint->setDistance(h[0].t);
    hitpoint->setPosition(ray.origin() + h[0].t * ray.direction());
    hitpoint->setNormal(_tris[h[0].tri].normal());
    return h[0].tri;
}

void QBVHAccel::intersect(const std::vector<Ray>& rays, std::vector<Intersection>* out) const {
    if (!_built) throw Error(PPPM_ERR_STATE, "QBVHAccel::construct not called");
    std::vector<pppm_hit> h;
    batch_intersect(_dev->handle(), *_dev, rays, &h);
    fill_hits(h, rays, _tris, out);
}

// ---- Scene -----------------------------------------------------------------
Scene::Scene() : _dev(Device::shared()) {}
Scene::Scene(std::shared_ptr<Device> dev) : _dev(std::move(dev)) {}

void Scene::addAreaLight(const Vector3D& c, const Vector3D& eu, const Vector3D& ev, const Vector3D& le) {
    pppm_bsdf b;
    std::memset(&b, 0, sizeof(b));
    b.type = PPPM_BSDF_EMITTER;
    b.emission[0] = (float)le.x();
    b.emission[1] = (float)le.y();
    b.emission[2] = (float)le.z();
    // one emitter material for the two triangles; their normal is eu x ev
    const int id = (int)_bsdfs.size();
    _bsdfs.push_back(BSDF(b));
    _triangles.push_back(Triangle(c, c + eu, c + eu + ev));
    _triangles.push_back(Triangle(c, c + eu + ev, c + ev));
    _bsdfIds.insert(_bsdfIds.end(), 2, id);
    pppm_light l;
    for (int d = 0; d < 3; d++) {
        l.corner[d] = (float)c[d];
        l.edge_u[d] = (float)eu[d];
        l.edge_v[d] = (float)ev[d];
        l.emission[d] = (float)le[d];
    }
    _lights.push_back(l);
    _accel = false;
}

PerspectiveCamera Scene::loadProcedural(int kind, uint64_t seed, int frequency, int width, int height) {
    pppm_scene_desc d;
    int rc = pppm_scenegen(kind, seed, frequency, width, height, &d);
    if (rc != PPPM_OK) throw Error(rc, "pppm_scenegen failed");
    clear();
    for (int32_t i = 0; i < d.n_bsdf; i++) _bsdfs.push_back(BSDF(d.bsdfs[i]));
    _triangles.reserve((size_t)d.n_tris);
    for (int64_t i = 0; i < d.n_tris; i++) {
        const float* t = d.tri_xyz + 9 * i;
        _triangles.push_back(Triangle(Vector3D(t[0], t[1], t[2]), Vector3D(t[3], t[4], t[5]), Vector3D(t[6], t[7], t[8])));
        _bsdfIds.push_back(d.bsdf_id[i]);
    }
    _lights.assign(d.lights, d.lights + d.n_lights);
    const pppm_camera& c = d.camera;
    PerspectiveCamera cam(Vector3D(c.center[0], c.center[1], c.center[2]),
                          Vector3D(c.direction[0], c.direction[1], c.direction[2]),
                          Vector3D(c.up[0], c.up[1], c.up[2]), c.fov_y_deg, c.width, c.height, 1.0);
    pppm_scenegen_free(&d);
    return cam;
}

void Scene::clear() {
    _triangles.clear();
    _bsdfIds.clear();
    _bsdfs.clear();
    _lights.clear();
    _accel = false;
}

void Scene::upload() const {
    std::vector<float> v = flatten(_triangles);
    std::vector<pppm_bsdf> table;
    for (const BSDF& b : _bsdfs) table.push_back(b.raw());
    _dev->check(pppm_scene_upload(_dev->handle(), v.data(), _bsdfIds.data(), (int64_t)_triangles.size(),
                                  table.data(), (int32_t)table.size(), _lights.data(), (int32_t)_lights.size()));
    _dev->check(pppm_bvh_build(_dev->handle(), nullptr));
}

void Scene::setAccelerator() {
    upload();
    _accel = true;
}

bool Scene::intersect(const Ray& ray, Intersection& isect) const {
    if (!_accel) throw Error(PPPM_ERR_STATE, "Scene::setAccelerator not called");
    std::vector<pppm_hit> h;
    batch_intersect(_dev->handle(), *_dev, std::vector<Ray>(1, ray), &h);
    std::vector<Intersection> out;
    fill_hits(h, std::vector<Ray>(1, ray), _triangles, &out);
    isect = out[0];
    return h[0].tri >= 0;
}

BBox Scene::bounds() const {
    BBox b;
    for (const Triangle& t : _triangles) b.merge(t);
    return b;
}

// ---- ProgressivePhotonMappingProb ------------------------------------------
const double ProgressivePhotonMappingProb::ALPHA = 0.7;  // ppm_probabilistic.cc:9

void ProgressivePhotonMappingProb::render(const Scene& scene, const Camera& camera,
                                          const RenderParameters& params) {
    std::shared_ptr<Device> dev = scene.device();
    void* ctx = dev->handle();
    scene.upload();  // the scene on this device, BVH included
    dev->check(pppm_camera_set(ctx, &camera.raw()));
    BBox b = scene.bounds();
    pppm_render_params p;
    std::memset(&p, 0, sizeof(p));
    p.photons = params.photons();
    p.passes = params.spp();
    p.max_bounces = _settings.bounceLimit;
    p.alpha = _settings.alpha;
    p.radius1 = _settings.radiusFraction * (b.posMax() - b.posMin()).norm();
    p.ray_eps = _settings.rayEps;
    p.seed = _settings.seed;
    p.scene_id = _settings.sceneId;
    p.hash_size = _settings.hashSize;
    p.capacity_factor = 8.0;
    p.gather_mode = _settings.gatherMode;
    dev->check(pppm_subsurface_configure(ctx, _settings.subsurfacePoints));
    dev->check(pppm_camera_set(ctx, &camera.raw()));  // configure ends a running render
    dev->check(pppm_render_begin(ctx, &p));
    if (_settings.learnedProposal) dev->check(pppm_proposal_enable(ctx, &_settings.proposal));
    _result = Image(camera.width(), camera.height());
    _stats.clear();
    for (int t = 1; t <= params.spp(); t++) {
        pppm_pass_stats st;
        dev->check(pppm_render_pass(ctx, &st));
        _stats.push_back(st);
        if (_onPass) {
            dev->check(pppm_image_read(ctx, _result.data().data(), (int64_t)_result.data().size()));
            _onPass(t, _result);
        }
    }
    dev->check(pppm_image_read(ctx, _result.data().data(), (int64_t)_result.data().size()));
}

}  // namespace pppm_amd
