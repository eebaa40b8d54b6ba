// pppm_amd.cc -- the reference-shaped C++ classes over the C ABI.
#include "pppm_amd.h"

#include <algorithm>
#include <cctype>
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
BSDF PhongBRDF::factory(const Vector3D& r, double n) {
    (void)n;
    return make_bsdf(PPPM_BSDF_PHONG, r);
}
BSDF SpecularBRDF::factory(const Vector3D& r) { return make_bsdf(PPPM_BSDF_SPECULAR, r); }
BSDF RefractionBSDF::factory(const Vector3D& r) { return make_bsdf(PPPM_BSDF_REFRACTION, r); }

PerspectiveCamera::PerspectiveCamera() { std::memset(&_c, 0, sizeof(_c)); }

Ray PerspectiveCamera::getRay(double pixelX, double pixelY) const {
    const Vector3D D = direction().normalized();
    const Vector3D U = Vector3D::cross(D, Vector3D(_c.up[0], _c.up[1], _c.up[2])).normalized();
    const Vector3D V = Vector3D::cross(U, D);
    const double th = std::tan(_c.fov_y_deg * 0.5 * M_PI / 180.0);
    const double aspect = (double)_c.width / (double)_c.height;
    const Vector3D dir = D + U * (th * aspect * (2.0 * pixelX / _c.width - 1.0)) +
                         V * (th * (2.0 * pixelY / _c.height - 1.0));
    return Ray(center(), dir.normalized());
}

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

void Image::setPixel(int x, int y, const Vector3D& c) {
    const size_t k = 3 * ((size_t)y * _w + x);
    _rgb[k] = c.x();
    _rgb[k + 1] = c.y();
    _rgb[k + 2] = c.z();
}

void Image::resize(int w, int h) {
    _w = w;
    _h = h;
    _rgb.assign(3 * (size_t)w * h, 0.0);
}

void Image::fill(const Vector3D& c) {
    for (size_t i = 0; i < _rgb.size(); i += 3) {
        _rgb[i] = c.x();
        _rgb[i + 1] = c.y();
        _rgb[i + 2] = c.z();
    }
}

void Image::gamma(double gam, bool inv) {
    const double e = inv ? 1.0 / gam : gam;
    for (double& v : _rgb) v = std::pow(v, e);
}

void Image::tonemap() {
    const size_t n = (size_t)_w * _h;
    if (n == 0) return;
    const double key = 0.18, tiny = 1.0e-8;
    double log_sum = 0.0, white = 0.0;
    for (size_t i = 0; i < n; i++) {
        const double lum = 0.2126 * _rgb[3 * i] + 0.7152 * _rgb[3 * i + 1] + 0.0722 * _rgb[3 * i + 2];
        log_sum += std::log(lum + tiny);
        white = std::max(white, lum);
    }
    const double mean_log = std::exp(log_sum / (double)n);
    const double white2 = white * white;
    for (double& v : _rgb) {
        const double s = v * key / mean_log;
        v = s * (1.0 + s / white2) / (1.0 + s);
    }
}

namespace {
std::string lower_ext(const std::string& f) {
    const size_t dot = f.find_last_of('.');
    std::string e = dot == std::string::npos ? "" : f.substr(dot);
    for (char& c : e) c = (char)std::tolower((unsigned char)c);
    return e;
}
}  // namespace

void Image::load(const std::string& filename) {
    double* rgb = nullptr;
    int32_t w = 0, h = 0;
    char err[512] = {0};
    const int rc = pppm_image_load(filename.c_str(), &rgb, &w, &h, err, (int32_t)sizeof(err));
    if (rc != PPPM_OK) throw Error(rc, std::string("Image::load: ") + err);
    _w = w;
    _h = h;
    _rgb.assign(rgb, rgb + 3 * (size_t)w * h);
    pppm_image_free(rgb);
}

void Image::save(const std::string& filename) const {
    const int rc = pppm_image_write(filename.c_str(), _rgb.data(), _w, _h, 0.0);
    if (rc != PPPM_OK) throw Error(rc, "Image::save: cannot write " + filename);
}

void Image::loadBmp(const std::string& filename) {
    if (lower_ext(filename) != ".bmp") throw Error(PPPM_ERR_INVALID, "loadBmp: not a .bmp file");
    load(filename);
}

void Image::saveBmp(const std::string& filename) const {
    if (lower_ext(filename) != ".bmp") throw Error(PPPM_ERR_INVALID, "saveBmp: not a .bmp file");
    save(filename);
}

void Image::loadHdr(const std::string& filename) {
    if (lower_ext(filename) != ".hdr") throw Error(PPPM_ERR_INVALID, "loadHdr: not a .hdr file");
    load(filename);
}

void Image::saveHdr(const std::string& filename) const {
    if (lower_ext(filename) != ".hdr") throw Error(PPPM_ERR_INVALID, "saveHdr: not a .hdr file");
    save(filename);
}

// ---- Envmap ----------------------------------------------------------------
Envmap::Envmap(const std::string& filename) {
    double* rgb = nullptr;
    int32_t w = 0, h = 0;
    char err[512] = {0};
    const int rc = pppm_image_load(filename.c_str(), &rgb, &w, &h, err, (int32_t)sizeof(err));
    if (rc != PPPM_OK) throw Error(rc, std::string("Envmap: ") + err);
    _w = w;
    _h = h;
    _rgb.resize(3 * (size_t)w * h);
    for (size_t i = 0; i < _rgb.size(); i++) _rgb[i] = (float)rgb[i];
    pppm_image_free(rgb);
}

void Envmap::resize(int width, int height) {
    _w = width;
    _h = height;
    _rgb.assign(3 * (size_t)width * height, 0.0f);
}

void Envmap::clearColor(const Vector3D& c) {
    for (size_t i = 0; i < _rgb.size(); i += 3) {
        _rgb[i] = (float)c.x();
        _rgb[i + 1] = (float)c.y();
        _rgb[i + 2] = (float)c.z();
    }
}

Vector3D Envmap::sampleFromDir(const Vector3D& d) const {
    if (empty()) return Vector3D();
    const double theta = std::acos(d.y());
    double phi = std::atan2(d.z(), d.x());
    if (phi < 0.0) phi += 2.0 * M_PI;
    const int ix = std::min(std::max((int)(phi / (2.0 * M_PI) * _w), 0), _w - 1);
    const int iy = std::min(std::max((int)(theta / M_PI * _h), 0), _h - 1);
    const float* p = &_rgb[3 * ((size_t)iy * _w + ix)];
    return Vector3D(p[0], p[1], p[2]);
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

// ---- Trimesh (sources/trimesh.cc) -------------------------------------------
Trimesh::Trimesh(const std::vector<Vector3D>& vertices, const std::vector<Triplet>& faces)
    : _vertices(vertices), _faces(faces) {
    _normals.resize(_faces.size());
    for (size_t i = 0; i < _faces.size(); i++) {
        const Vector3D& v0 = _vertices.at((size_t)_faces[i][0]);
        _normals[i] = Vector3D::cross(_vertices.at((size_t)_faces[i][1]) - v0, _vertices.at((size_t)_faces[i][2]) - v0)
                          .normalized();
    }
}

void Trimesh::load(const std::string& filename) {
    pppm_mesh m;
    char err[512] = {0};
    const int rc = pppm_mesh_load(filename.c_str(), &m, err, (int32_t)sizeof(err));
    if (rc != PPPM_OK) throw Error(rc, std::string("Trimesh::load: ") + err);
    _vertices.resize((size_t)m.n_vertices);
    for (int64_t i = 0; i < m.n_vertices; i++)
        _vertices[(size_t)i] = Vector3D(m.vertices[3 * i], m.vertices[3 * i + 1], m.vertices[3 * i + 2]);
    _faces.resize((size_t)m.n_faces);
    _normals.resize((size_t)m.n_faces);
    for (int64_t i = 0; i < m.n_faces; i++) {
        const int32_t* f = m.faces + 3 * i;
        _faces[(size_t)i] = Triplet(f[0], f[1], f[2]);
        _normals[(size_t)i] =
            Vector3D::cross(_vertices[(size_t)f[1]] - _vertices[(size_t)f[0]], _vertices[(size_t)f[2]] - _vertices[(size_t)f[0]])
                .normalized();
    }
    pppm_mesh_free(&m);
    _accel.reset();
}

void Trimesh::translate(const Vector3D& move) {
    for (Vector3D& v : _vertices) v += move;
}

void Trimesh::scale(double sx, double sy, double sz) {
    for (Vector3D& v : _vertices) v = Vector3D(v.x() * sx, v.y() * sy, v.z() * sz);
}

void Trimesh::scale(double s) { scale(s, s, s); }

void Trimesh::putOnPlane(const Plane& plane) {
    // the lowest vertex along the normal moves onto the plane (trimesh.cc:283-293)
    double lowest = 1.0e20;
    for (const Vector3D& v : _vertices) lowest = std::min(lowest, Vector3D::dot(plane.normal(), v));
    for (Vector3D& v : _vertices) v -= (lowest + plane.distance()) * plane.normal();
}

void Trimesh::fitToBBox(const BBox& bbox) {
    // uniform scale to the tightest axis, then centre to centre (trimesh.cc:295-312)
    BBox org;
    for (const Vector3D& v : _vertices) org.merge(v);
    const Vector3D target = bbox.posMax() - bbox.posMin();
    const Vector3D size = org.posMax() - org.posMin();
    const double s = std::min(target.x() / size.x(), std::min(target.y() / size.y(), target.z() / size.z()));
    scale(s);
    const Vector3D from = (org.posMin() + org.posMax()) * (0.5 * s);
    const Vector3D to = (bbox.posMin() + bbox.posMax()) * 0.5;
    translate(to - from);
}

Triangle Trimesh::getTriangle(int id) const {
    const Triplet& f = _faces.at((size_t)id);
    return Triangle(_vertices[(size_t)f[0]], _vertices[(size_t)f[1]], _vertices[(size_t)f[2]]);
}

std::vector<Triangle> Trimesh::triangulate() const {
    std::vector<Triangle> t(_faces.size());
    for (size_t i = 0; i < _faces.size(); i++) t[i] = getTriangle((int)i);
    return t;
}

double Trimesh::area() const {
    double a = 0.0;
    for (size_t i = 0; i < _faces.size(); i++) a += getTriangle((int)i).area();
    return a;
}

void Trimesh::buildAccel() {
    _accel = std::make_shared<QBVHAccel>(Device::shared());
    _accel->construct(triangulate());
}

bool Trimesh::intersect(const Ray& ray, Hitpoint* hitpoint) const {
    if (!_accel) throw Error(PPPM_ERR_STATE, "Trimesh::buildAccel not called");
    hitpoint->setDistance(1.0e20);
    return _accel->intersect(ray, hitpoint) != -1;
}

// ---- HostBvh ---------------------------------------------------------------
void HostBvh::download(const Device& dev) {
    int64_t nn = 0, nt = 0;
    int32_t root = 0x7FFFFFFF;
    dev.check(pppm_bvh_download(dev.handle(), nullptr, 0, nullptr, 0, &nn, &nt, &root));
    _nodes.resize((size_t)nn);
    _tris.resize(12 * (size_t)nt);
    dev.check(pppm_bvh_download(dev.handle(), _nodes.data(), nn, _tris.data(), nt, &nn, &nt, &root));
    _root = root;
}

int HostBvh::intersect(const Ray& ray, double tmax, double* t_out) const {
    const double EPS = 1.0e-12;  // sources/common.h:14
    if (empty()) return -1;
    const Vector3D o = ray.origin(), d = ray.direction();
    float org[3], inv[3];
    for (int k = 0; k < 3; k++) {
        org[k] = (float)o[k];
        inv[k] = d[k] == 0.0 ? 1.0e20f : (float)(1.0 / d[k]);  // qbvh_accel.cc:211-233
    }
    double best = tmax;
    int hit = -1;
    int32_t stack[256];
    int sp = 0;
    stack[sp++] = _root;
    while (sp > 0) {
        const int32_t ref = stack[--sp];
        if (ref >= 0) {
            const pppm_bvh_node& n = _nodes[(size_t)ref];
            const float* lo[3] = {n.lox, n.loy, n.loz};
            const float* hi[3] = {n.hix, n.hiy, n.hiz};
            float tn[4];
            int order[4], m = 0;
            for (int c = 0; c < 4; c++) {
                if (n.child[c] == 0x7FFFFFFF) continue;
                float a = 0.0f, b = (float)best;
                for (int k = 0; k < 3; k++) {
                    float t0 = (lo[k][c] - org[k]) * inv[k], t1 = (hi[k][c] - org[k]) * inv[k];
                    a = std::fmax(a, std::fmin(t0, t1));
                    b = std::fmin(b, std::fmax(t0, t1) * 1.00000024f);  // conservative exit (Ize 2013)
                }
                if (a <= b) {
                    tn[c] = a;
                    order[m++] = c;
                }
            }
            // push far to near: the nearest child is popped first
            for (int i = 1; i < m; i++)
                for (int j = i; j > 0 && tn[order[j]] > tn[order[j - 1]]; j--) std::swap(order[j], order[j - 1]);
            for (int i = 0; i < m && sp < 256; i++) stack[sp++] = n.child[order[i]];
        } else {
            const int32_t e = ~ref;
            const int32_t first = e >> 3, count = (e & 7) + 1;
            for (int k = 0; k < count; k++) {
                const float* t = &_tris[12 * (size_t)(first + k)];
                const Vector3D p0(t[0], t[1], t[2]), p1(t[4], t[5], t[6]), p2(t[8], t[9], t[10]);
                // sources/triangle.cc:53-78 (Moller-Trumbore, FP64)
                const Vector3D e1 = p1 - p0, e2 = p2 - p0;
                const Vector3D pv = Vector3D::cross(d, e2);
                const double det = Vector3D::dot(e1, pv);
                if (det > -EPS && det < EPS) continue;
                const double invdet = 1.0 / det;
                const Vector3D tv = o - p0;
                const double u = Vector3D::dot(tv, pv) * invdet;
                if (u < 0.0 || u > 1.0) continue;
                const Vector3D qv = Vector3D::cross(tv, e1);
                const double v = Vector3D::dot(d, qv) * invdet;
                if (v < 0.0 || u + v > 1.0) continue;
                const double tt = Vector3D::dot(e2, qv) * invdet;
                if (tt > EPS && tt < best) {
                    best = tt;
                    int32_t id;
                    std::memcpy(&id, &t[3], 4);
                    hit = id;
                }
            }
        }
    }
    if (t_out) *t_out = best;
    return hit;
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
    _host.download(*_dev);
    _built = true;
}

int QBVHAccel::intersect(const Ray& ray, Hitpoint* hitpoint) const {
    if (!_built) throw Error(PPPM_ERR_STATE, "QBVHAccel::construct not called");
    double t = 0.0;
    const int id = _host.intersect(ray, hitpoint->distance(), &t);
    if (id < 0) return -1;
    hitpoint->setDistance(t);
    hitpoint->setPosition(ray.origin() + t * ray.direction());
    hitpoint->setNormal(_tris[(size_t)id].normal());
    return id;
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
    if (!_envmap.empty())
        _dev->check(pppm_envmap_set(_dev->handle(), _envmap.texels().data(), _envmap.width(), _envmap.height()));
}

void Scene::setAccelerator() {
    upload();
    _host.download(*_dev);
    _accel = true;
}

bool Scene::intersect(const Ray& ray, Intersection& isect) const {
    if (!_accel) throw Error(PPPM_ERR_STATE, "Scene::setAccelerator not called");
    double t = 0.0;
    const int id = _host.intersect(ray, 1.0e20, &t);
    Hitpoint hp;
    if (id >= 0) {
        hp.setDistance(t);
        hp.setPosition(ray.origin() + t * ray.direction());
        hp.setNormal(_triangles[(size_t)id].normal());
    }
    isect.setHitpoint(hp);
    isect.setObjectId(id);
    return id >= 0;
}

BBox Scene::bounds() const {
    BBox b;
    for (const Triangle& t : _triangles) b.merge(t);
    return b;
}

// ---- ProgressivePhotonMappingProb ------------------------------------------
const double ProgressivePhotonMappingProb::ALPHA = 0.7;  // ppm_probabilistic.cc:9

namespace {

// The render loop shared by both methods: params.spp() passes of
// params.photons() photons each.
void run_passes(const Scene& scene, const Camera& camera, const RenderParameters& params, const PPMSettings& set,
                int32_t method, Image& result, std::vector<pppm_pass_stats>& stats,
                const std::function<void(int, const Image&)>& onPass) {
    if (camera.width() <= 0 || camera.height() <= 0)
        throw Error(PPPM_ERR_INVALID, "the camera has no image: assign a PerspectiveCamera with a size");
    std::shared_ptr<Device> dev = scene.device();
    void* ctx = dev->handle();
    scene.upload();  // the scene on this device, BVH included
    dev->check(pppm_camera_set(ctx, &camera.raw()));
    BBox b = scene.bounds();
    pppm_render_params p;
    std::memset(&p, 0, sizeof(p));
    p.photons = params.photons();
    p.passes = params.spp();
    p.max_bounces = set.bounceLimit;
    p.alpha = set.alpha;
    p.radius1 = set.radiusFraction * (b.posMax() - b.posMin()).norm();
    p.ray_eps = set.rayEps;
    p.seed = set.seed;
    p.scene_id = set.sceneId;
    p.hash_size = set.hashSize;
    p.capacity_factor = 8.0;
    p.gather_mode = set.gatherMode;
    p.method = method;
    if (method == PPPM_METHOD_PROBABILISTIC) {
        dev->check(pppm_subsurface_configure(ctx, set.subsurfacePoints));
        dev->check(pppm_camera_set(ctx, &camera.raw()));  // configure ends a running render
    }
    dev->check(pppm_render_begin(ctx, &p));
    if (method == PPPM_METHOD_PROBABILISTIC && set.learnedProposal) dev->check(pppm_proposal_enable(ctx, &set.proposal));
    result = Image(camera.width(), camera.height());
    stats.clear();
    for (int t = 1; t <= params.spp(); t++) {
        pppm_pass_stats st;
        dev->check(pppm_render_pass(ctx, &st));
        stats.push_back(st);
        if (onPass) {
            dev->check(pppm_image_read(ctx, result.data().data(), (int64_t)result.data().size()));
            onPass(t, result);
        }
    }
    dev->check(pppm_image_read(ctx, result.data().data(), (int64_t)result.data().size()));
}

}  // namespace

void ProgressivePhotonMappingProb::render(const Scene& scene, const Camera& camera,
                                          const RenderParameters& params, RandomSamplerType randomSamplerType) {
    if (randomSamplerType != RANDOM_SAMPLER_PSEUDO_RANDOM)
        throw Error(PPPM_ERR_INVALID, "only RANDOM_SAMPLER_PSEUDO_RANDOM is provided: the GPU path draws counter-based "
                                      "Philox numbers and has no permuted Halton sequence");
    run_passes(scene, camera, params, _settings, PPPM_METHOD_PROBABILISTIC, _result, _stats, _onPass);
}

const double ProgressivePhotonMapping::ALPHA = 0.7;  // progressive_photon_mapping.cc:17

void ProgressivePhotonMapping::render(const Scene& scene, const Camera& camera, const RenderParameters& params) {
    if (_settings.learnedProposal)
        throw Error(PPPM_ERR_INVALID, "the learned proposal belongs to the probabilistic method");
    run_passes(scene, camera, params, _settings, PPPM_METHOD_HACHISUKA, _result, _stats, _onPass);
}

}  // namespace pppm_amd
