// pppm_amd.h -- C++ interface of the MI355X photon mapper with the class
// names, signatures and call pattern of tatsy/tatsy-pppm (paths relative to
// the reference root), implemented over the C ABI of include/pppm.h.
//
//   Scene scene;                                    // sources/scene.h:26-67
//   scene.add(Triangle(a, b, c), LambertianBRDF::factory(rho));
//   scene.addAreaLight(corner, edgeU, edgeV, Le);   // new: the configs' emitter
//   scene.setAccelerator();                         // GPU 4-wide BVH build
//   PerspectiveCamera camera(center, dir, up, fovDeg, w, h, 1.0);
//   ProgressivePhotonMappingProb ppm;               // sources/ppm_probabilistic.h:23-42
//   ppm.render(scene, camera, RenderParameters(photons, passes, 0, 0.0));
//   const Image& img = ppm.result();
//
// Everything lives in namespace pppm_amd; `using namespace pppm_amd;` makes a
// reference call site compile unchanged.  Errors throw pppm_amd::Error (the
// reference aborts through Assertion, sources/common.h:24-35).
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/pppm.h"

namespace pppm_amd {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// ---- sources/vector3d.h:21-98 (FP64 vector) --------------------------------
class Vector3D {
    double _x, _y, _z;

public:
    Vector3D() : _x(0.0), _y(0.0), _z(0.0) {}
    Vector3D(double x, double y, double z) : _x(x), _y(y), _z(z) {}
    double x() const { return _x; }
    double y() const { return _y; }
    double z() const { return _z; }
    double operator[](int d) const { return d == 0 ? _x : (d == 1 ? _y : _z); }
    static double dot(const Vector3D& u, const Vector3D& v);
    static Vector3D cross(const Vector3D& u, const Vector3D& v);
    static Vector3D minimum(const Vector3D& u, const Vector3D& v);
    static Vector3D maximum(const Vector3D& u, const Vector3D& v);
    double norm() const;
    double squaredNorm() const { return dot(*this, *this); }
    Vector3D normalized() const;
    Vector3D operator-() const { return Vector3D(-_x, -_y, -_z); }
    Vector3D& operator+=(const Vector3D& v);
    Vector3D& operator-=(const Vector3D& v);
    Vector3D& operator*=(double s);
};
Vector3D operator+(const Vector3D& u, const Vector3D& v);
Vector3D operator-(const Vector3D& u, const Vector3D& v);
Vector3D operator*(const Vector3D& u, const Vector3D& v);
Vector3D operator*(const Vector3D& v, double s);
Vector3D operator*(double s, const Vector3D& v);
Vector3D operator/(const Vector3D& v, double s);

// ---- sources/ray.h:16-79 ---------------------------------------------------
class Ray {
    Vector3D _origin, _direction;

public:
    Ray() {}
    Ray(const Vector3D& origin, const Vector3D& direction) : _origin(origin), _direction(direction) {}
    Vector3D origin() const { return _origin; }
    Vector3D direction() const { return _direction; }
};

class Hitpoint {
    double _distance = 1.0e20;  // INFTY, sources/ray.cc:48-52
    Vector3D _normal, _position;

public:
    double distance() const { return _distance; }
    Vector3D normal() const { return _normal; }
    Vector3D position() const { return _position; }
    void setDistance(double d) { _distance = d; }
    void setNormal(const Vector3D& n) { _normal = n; }
    void setPosition(const Vector3D& p) { _position = p; }
};

class Intersection {
    Hitpoint _hitPoint;
    int _objectId = -1;

public:
    int objectID() const { return _objectId; }
    const Hitpoint& hitpoint() const { return _hitPoint; }
    double hittingDistance() const { return _hitPoint.distance(); }
    void setHitpoint(const Hitpoint& h) { _hitPoint = h; }
    void setObjectId(int id) { _objectId = id; }
};

class Triangle;

// ---- sources/bbox.h:21-56 --------------------------------------------------
class BBox {
    Vector3D _posMin, _posMax;

public:
    BBox() : _posMin(1e20, 1e20, 1e20), _posMax(-1e20, -1e20, -1e20) {}
    BBox(const Vector3D& mn, const Vector3D& mx) : _posMin(mn), _posMax(mx) {}
    const Vector3D& posMin() const { return _posMin; }
    const Vector3D& posMax() const { return _posMax; }
    void merge(const Vector3D& v);
    void merge(const BBox& b);
    void merge(const Triangle& t);
};

// ---- sources/triangle.h:16-45 and geometry_interface.h ----------------------
class IGeometry {
public:
    virtual ~IGeometry() {}
    virtual std::vector<Triangle> triangulate() const = 0;
};

class Triangle : public IGeometry {
    Vector3D _p0, _p1, _p2;

public:
    Triangle() {}
    Triangle(const Vector3D& p0, const Vector3D& p1, const Vector3D& p2) : _p0(p0), _p1(p1), _p2(p2) {}
    Vector3D p(int id) const { return id == 0 ? _p0 : (id == 1 ? _p1 : _p2); }
    Vector3D p0() const { return _p0; }
    Vector3D p1() const { return _p1; }
    Vector3D p2() const { return _p2; }
    Vector3D gravity() const { return (_p0 + _p1 + _p2) / 3.0; }
    Vector3D normal() const { return Vector3D::cross(_p1 - _p0, _p2 - _p0).normalized(); }
    double area() const { return 0.5 * Vector3D::cross(_p1 - _p0, _p2 - _p1).norm(); }
    std::vector<Triangle> triangulate() const override { return std::vector<Triangle>(1, *this); }
};

// ---- sources/plane.h:16-41 ---------------------------------------------------
// The plane dot(n, x) + distance = 0 (the reference's putOnPlane convention);
// a placement target, not renderable geometry.
class Plane {
    double _distance = 0.0;
    Vector3D _normal;

public:
    Plane() {}
    Plane(double distance, const Vector3D& normal) : _distance(distance), _normal(normal) {}
    double distance() const { return _distance; }
    Vector3D normal() const { return _normal; }
};

// ---- sources/trimesh.h:19-140 ------------------------------------------------
class Triplet {
    int _d[3] = {0, 0, 0};

public:
    Triplet() {}
    Triplet(int i, int j, int k) : _d{i, j, k} {}
    int operator[](int i) const { return _d[i]; }
};

class QBVHAccel;

// Indexed triangle mesh in FP64 with the reference's placement transforms
// (translate, scale, putOnPlane, fitToBBox: the same arithmetic in the same
// order, so the same triangles bit for bit).  load() reads .ply / .obj
// through pppm_mesh_load.
class Trimesh : public IGeometry {
    std::vector<Vector3D> _vertices;
    std::vector<Vector3D> _normals;  // per face
    std::vector<Triplet> _faces;
    std::shared_ptr<QBVHAccel> _accel;

public:
    Trimesh() {}
    explicit Trimesh(const std::string& filename) { load(filename); }
    Trimesh(const std::vector<Vector3D>& vertices, const std::vector<Triplet>& faces);
    void load(const std::string& filename);
    void translate(const Vector3D& move);
    void scale(double scaleX, double scaleY, double scaleZ);
    void scale(double scaleAll);
    void putOnPlane(const Plane& plane);
    void fitToBBox(const class BBox& bbox);
    std::vector<Triangle> triangulate() const override;
    double area() const;
    size_t numVerts() const { return _vertices.size(); }
    size_t numFaces() const { return _faces.size(); }
    std::vector<Triplet> getIndices() const { return _faces; }
    Vector3D getVertex(int id) const { return _vertices.at((size_t)id); }
    Vector3D getNormal(int id) const { return _normals.at((size_t)id); }
    Triangle getTriangle(int id) const;
    // GPU BVH over the mesh alone, and a ray against it (trimesh.cc:89-94, 117-129)
    void buildAccel();
    bool intersect(const Ray& ray, Hitpoint* hitpoint) const;
};

// ---- sources/bsdf.h:19-59, sources/brdf.h:37-101 ---------------------------
enum BsdfType {
    BSDF_TYPE_NONE = 0x00,
    BSDF_TYPE_LAMBERTIAN_BRDF = PPPM_BSDF_LAMBERTIAN,
    BSDF_TYPE_SPECULAR_BRDF = PPPM_BSDF_SPECULAR,
    BSDF_TYPE_PHONG_BRDF = PPPM_BSDF_PHONG,
    BSDF_TYPE_REFRACTION = PPPM_BSDF_REFRACTION,
    BSDF_TYPE_BSSRDF = PPPM_BSDF_BSSRDF,
    BSDF_TYPE_EMITTER = PPPM_BSDF_EMITTER
};

// ---- sources/bssrdf.h:17-76 (DipoleBSSRDF) ---------------------------------
// The dipole parameters; scale must be 1 (the reference's only use,
// sources/main/appmain.cc:64).
class BSSRDF {
    Vector3D _sigma_a, _sigmap_s;
    double _eta = 1.3;

public:
    BSSRDF() {}
    BSSRDF(const Vector3D& sigma_a, const Vector3D& sigmap_s, double eta)
        : _sigma_a(sigma_a), _sigmap_s(sigmap_s), _eta(eta) {}
    const Vector3D& sigma_a() const { return _sigma_a; }
    const Vector3D& sigmap_s() const { return _sigmap_s; }
    double eta() const { return _eta; }
};
struct DipoleBSSRDF {
    static BSSRDF factory(const Vector3D& sigma_a, const Vector3D& sigmap_s, double eta = 1.3, double scale = 1.0);
};

// Value type holding the material parameters the GPU table needs.
class BSDF {
    pppm_bsdf _b;

public:
    BSDF();
    explicit BSDF(const pppm_bsdf& b) : _b(b) {}
    Vector3D reflectance() const;
    BsdfType type() const { return (BsdfType)_b.type; }
    const pppm_bsdf& raw() const { return _b; }
    // sources/bsdf.cc:70-73: attaches a BSSRDF and adds BSDF_TYPE_BSSRDF
    void setBssrdf(const BSSRDF& bssrdf);
};

struct LambertianBRDF {
    static BSDF factory(const Vector3D& reflectance);
};
// sources/brdf.h:75-101.  Phong lobes have no GPU kernel (no workload uses
// them): a scene holding one is rejected by Scene::setAccelerator.
struct PhongBRDF {
    static BSDF factory(const Vector3D& reflectance, double n = 128.0);
};
struct SpecularBRDF {
    static BSDF factory(const Vector3D& reflectance);
};
struct RefractionBSDF {
    static BSDF factory(const Vector3D& reflectance);
};

// ---- sources/size.h ---------------------------------------------------------
template <class T>
class Size {
    T _w = T(), _h = T();

public:
    Size() {}
    Size(T w, T h) : _w(w), _h(h) {}
    T width() const { return _w; }
    T height() const { return _h; }
};

// ---- sources/perspective_camera.h:16-31, camera_interface.h:12-81 -----------
// fov is the vertical field of view in DEGREES (the unit the reference's own
// caller passes, sources/main/appmain.cc:123; its implementation reads
// radians, sources/perspective_camera.cc:18), and the image basis is
// orthonormal (camera_interface.h:40-41 leaves it unnormalised).  The
// default camera has an empty image; assign a real one before rendering, as
// the reference's application does (appmain.cc:17, :123).
class PerspectiveCamera {
    pppm_camera _c;
    double _sensitivity = 0.0;

public:
    PerspectiveCamera();
    PerspectiveCamera(const Vector3D& center, const Vector3D& direction, const Vector3D& up,
                      double fov, int imageWidth, int imageHeight, double sensitivity);
    // camera ray through image position (pixelX, pixelY), pixelY upward from
    // the bottom edge (perspective_camera.cc:42-46): origin at the centre,
    // unit direction
    Ray getRay(double pixelX, double pixelY) const;
    const pppm_camera& raw() const { return _c; }
    int width() const { return _c.width; }
    int height() const { return _c.height; }
    Size<int> imagesize() const { return Size<int>(_c.width, _c.height); }
    Vector3D center() const { return Vector3D(_c.center[0], _c.center[1], _c.center[2]); }
    Vector3D direction() const { return Vector3D(_c.direction[0], _c.direction[1], _c.direction[2]); }
    double sensitivity() const { return _sensitivity; }
};
typedef PerspectiveCamera Camera;

// ---- sources/random_sampler.h:21-24 ------------------------------------------
// The GPU path draws counter-based Philox-4x32-10 numbers keyed by (seed,
// pass, path, dimension) whichever sampler is named; the permuted Halton
// sequence (quasi-Monte Carlo) is not provided and is rejected by render().
enum RandomSamplerType { RANDOM_SAMPLER_PSEUDO_RANDOM, RANDOM_SAMPLER_QUASI_MONTE_CARLO };

// ---- sources/render_parameters.h:4-44 --------------------------------------
class RenderParameters {
    int _photons, _spp, _gatherPhotons;
    double _gatherRadius;

public:
    explicit RenderParameters(int photons = 1000000, int spp = 16, int gatherPhotons = 16,
                              double gatherRadius = 16.0)
        : _photons(photons), _spp(spp), _gatherPhotons(gatherPhotons), _gatherRadius(gatherRadius) {}
    int photons() const { return _photons; }
    int spp() const { return _spp; }
    int gatherPhotons() const { return _gatherPhotons; }
    double gatherRadius() const { return _gatherRadius; }
};

// Knobs the reference hard-codes (ALPHA, ppm_probabilistic.cc:9; the initial
// radius, :55) and the extensions of the GPU path.
struct PPMSettings {
    double alpha = 0.7;
    double radiusFraction = 0.01;  // r_1 = fraction * scene bbox diagonal
    double rayEps = 1.0e-4;
    uint64_t seed = 0x7A75;
    uint32_t sceneId = 0;
    int bounceLimit = 64;
    int hashSize = 0;
    int gatherMode = PPPM_GATHER_AUTO;
    int64_t subsurfacePoints = 0;     // BSSRDF surface points (0: the reference's density, <= 2^20)
    bool learnedProposal = false;     // the learned emission mixture (one light)
    pppm_proposal_params proposal = {0.5, 2.0e-3, 16 << 20, 8, 8, 64, 16384, 0x5EED};
};

// ---- sources/image.h (FP64 RGB buffer) --------------------------------------
class Image {
    int _w = 0, _h = 0;
    std::vector<double> _rgb;

public:
    Image() {}
    Image(int w, int h) : _w(w), _h(h), _rgb(3 * (size_t)w * h, 0.0) {}
    int width() const { return _w; }
    int height() const { return _h; }
    Vector3D operator()(int x, int y) const;
    void setPixel(int x, int y, const Vector3D& c);
    std::vector<double>& data() { return _rgb; }
    const std::vector<double>& data() const { return _rgb; }
    void resize(int w, int h);
    void fill(const Vector3D& color);
    // every channel raised to gam (inv: to 1 / gam), image.cc:202-213
    void gamma(double gam, bool inv = false);
    // Reinhard's global operator with key 0.18 and the brightest luminance
    // as white, image.cc:498-524
    void tonemap();
    // by extension: .bmp, .png, .hdr (RGBE) and .pfm; the values as they are
    // (apply gamma() first), image.cc:215-240; errors throw
    void load(const std::string& filename);
    void save(const std::string& filename) const;
    void loadBmp(const std::string& filename);
    void saveBmp(const std::string& filename) const;
    void loadHdr(const std::string& filename);
    void saveHdr(const std::string& filename) const;
};

// ---- sources/envmap.h:14-60 ------------------------------------------------
// Lat-long environment (row 0 along +y).  A map of 64 x 64 texels or more
// lights the scene it is set on (Scene::setEnvmap) and is what eye rays
// leaving the scene see; the default (empty) map is black.
class Envmap {
    int _w = 0, _h = 0;
    std::vector<float> _rgb;  // [h][w][3]

public:
    Envmap() {}
    // .hdr (RGBE), .pfm or .bmp through pppm_image_load
    explicit Envmap(const std::string& filename);
    void resize(int width, int height);
    void clearColor(const Vector3D& color);
    // the texel of a direction (envmap.cc:47-60)
    Vector3D sampleFromDir(const Vector3D& dir) const;
    int width() const { return _w; }
    int height() const { return _h; }
    bool empty() const { return _rgb.empty(); }
    const std::vector<float>& texels() const { return _rgb; }
    std::vector<float>& texels() { return _rgb; }
};

// A device context shared by the objects that use the GPU.
class Device {
    void* _ctx = nullptr;

public:
    explicit Device(int device = 0);
    ~Device();
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;
    void* handle() const { return _ctx; }
    void check(int rc) const;
    static std::shared_ptr<Device> shared(int device = 0);
};

// Host copy of a GPU-built 4-wide BVH (pppm_bvh_download) and its scalar
// closest-hit traversal on the CPU: const, allocation-free and re-entrant, so
// any number of host threads may call it at once (the reference's OpenMP
// loops call QBVHAccel::intersect concurrently, ppm_probabilistic.cc:157, 256).
// Boxes are tested in FP32 with a conservative exit distance, triangles by
// the reference's FP64 Moller-Trumbore test (sources/triangle.cc:53-78).
class HostBvh {
    std::vector<pppm_bvh_node> _nodes;
    std::vector<float> _tris;  // [n][12] leaf order, v0.w = upload index bits
    int32_t _root = 0x7FFFFFFF;

public:
    void download(const Device& dev);
    bool empty() const { return _root == 0x7FFFFFFF; }
    // closest hit with t in (1e-12, tmax): upload index of the triangle, or -1
    int intersect(const Ray& ray, double tmax, double* t) const;
};

// ---- sources/qbvh_accel.h:13-62 --------------------------------------------
class QBVHAccel {
    std::shared_ptr<Device> _dev;
    std::vector<Triangle> _tris;
    HostBvh _host;
    bool _built = false;

public:
    QBVHAccel();
    explicit QBVHAccel(std::shared_ptr<Device> dev);
    // GPU build over the triangle soup (replaces the serial object-median build)
    void construct(const std::vector<Triangle>& triangles);
    // triangle id or -1; hitpoint->distance() is the ray's tMax on input.
    // Host traversal of the downloaded tree: re-entrant.
    int intersect(const Ray& ray, Hitpoint* hitpoint) const;
    // batch form on the GPU: the production path
    void intersect(const std::vector<Ray>& rays, std::vector<Intersection>* out) const;
};

// ---- sources/scene.h:26-67 -------------------------------------------------
class Scene {
    std::shared_ptr<Device> _dev;
    std::vector<Triangle> _triangles;
    std::vector<int> _bsdfIds;
    std::vector<BSDF> _bsdfs;
    std::vector<pppm_light> _lights;
    Envmap _envmap;
    HostBvh _host;
    bool _accel = false;

public:
    Scene();
    explicit Scene(std::shared_ptr<Device> dev);
    Scene(const Scene&) = delete;
    Scene& operator=(const Scene&) = delete;

    template <class Ty>
    void add(const Ty& primitive, const BSDF& bsdf) {
        static_assert(std::is_base_of<IGeometry, Ty>::value, "Ty must be an IGeometry");
        std::vector<Triangle> t = primitive.triangulate();
        _triangles.insert(_triangles.end(), t.begin(), t.end());
        _bsdfIds.insert(_bsdfIds.end(), t.size(), (int)_bsdfs.size());
        _bsdfs.push_back(bsdf);
    }
    // parallelogram emitter (two triangles + a photon source)
    void addAreaLight(const Vector3D& corner, const Vector3D& edgeU, const Vector3D& edgeV,
                      const Vector3D& emission);
    // load a procedural benchmark scene (include/pppm.h PPPM_SCENE_*)
    PerspectiveCamera loadProcedural(int kind, uint64_t seed, int frequency, int width, int height);

    // sources/scene.h:60-61: the environment lights the scene beside its area lights
    void setEnvmap(const Envmap& envmap) { _envmap = envmap; }
    const Envmap& envmap() const { return _envmap; }
    const Triangle& getTriangle(int id) const { return _triangles.at(id); }
    const BSDF& getBsdf(int triId) const { return _bsdfs.at(_bsdfIds.at(triId)); }
    size_t numTriangles() const { return _triangles.size(); }
    void clear();
    void setAccelerator();
    // host traversal of the tree setAccelerator built: const and re-entrant
    bool intersect(const Ray& ray, Intersection& isect) const;
    BBox bounds() const;
    std::shared_ptr<Device> device() const { return _dev; }
    void upload() const;
};

// ---- sources/hash_grid.h:8-39 ----------------------------------------------
// Same API.  add() records (item, box); the buckets are built on the GPU by a
// stable radix sort (construct(), or lazily on the first operator[]), with
// the bucket contents of the reference: every item of every cell its box
// covers, in insertion order.
template <class Ty>
class HashGrid {
    int _hashSize = -1;
    BBox _bbox;
    double _hashScale = 0.0;
    std::vector<Ty> _items;
    std::vector<double> _bmin, _bmax;
    std::vector<int32_t> _sorted, _start, _end;
    mutable std::vector<std::vector<Ty>> _data;
    bool _dirty = true;
    std::shared_ptr<Device> _dev;

public:
    HashGrid() : _dev(Device::shared()) {}
    void init(int hashSize, double hashScale, const BBox& bbox) {
        _hashSize = hashSize;
        _hashScale = hashScale;
        _bbox = bbox;
        clear();
    }
    void add(const Ty& p, const Vector3D& boxMin, const Vector3D& boxMax) {
        _items.push_back(p);
        _bmin.insert(_bmin.end(), {boxMin.x(), boxMin.y(), boxMin.z()});
        _bmax.insert(_bmax.end(), {boxMax.x(), boxMax.y(), boxMax.z()});
        _dirty = true;
    }
    void clear() {
        _items.clear();
        _bmin.clear();
        _bmax.clear();
        _data.clear();
        _dirty = true;
    }
    // builds the buckets on the GPU (the reference's construct is an empty stub)
    void construct(std::vector<Ty>& points = dummy(), const int imageW = -1, const int imageH = -1) {
        (void)points; (void)imageW; (void)imageH;
        if (_hashSize <= 0) throw Error(PPPM_ERR_STATE, "hash size is not initialized");
        const double bm[3] = {_bbox.posMin().x(), _bbox.posMin().y(), _bbox.posMin().z()};
        int64_t n = (int64_t)_items.size(), npairs = 0;
        _start.assign(_hashSize, 0);
        _end.assign(_hashSize, 0);
        int64_t cap = 8 * n + 16;
        for (int attempt = 0; attempt < 2; attempt++) {
            _sorted.assign((size_t)cap, 0);
            int rc = pppm_hashgrid_build_boxes(_dev->handle(), _bmin.data(), _bmax.data(), n, _hashSize,
                                               _hashScale, bm, _sorted.data(), cap, &npairs,
                                               _start.data(), _end.data());
            if (rc == PPPM_ERR_CAPACITY) { cap = npairs; continue; }
            _dev->check(rc);
            break;
        }
        _sorted.resize((size_t)npairs);
        _data.assign(_hashSize, std::vector<Ty>());
        for (int h = 0; h < _hashSize; h++)
            for (int32_t j = _start[h]; j < _end[h]; j++) _data[h].push_back(_items[_sorted[j]]);
        _dirty = false;
    }
    std::vector<Ty>& operator[](const Vector3D& v) {
        if (_dirty) construct();
        const double b[3] = {(v.x() - _bbox.posMin().x()) * _hashScale, (v.y() - _bbox.posMin().y()) * _hashScale,
                             (v.z() - _bbox.posMin().z()) * _hashScale};
        int c[3];
        for (int d = 0; d < 3; d++) {
            int i = static_cast<int>(b[d]);
            c[d] = i < 0 ? -i : i;
        }
        uint32_t h = ((uint32_t)c[0] * 73856093u) ^ ((uint32_t)c[1] * 19349663u) ^ ((uint32_t)c[2] * 83492791u);
        return _data[h % (uint32_t)_hashSize];
    }

private:
    static std::vector<Ty>& dummy() {
        static std::vector<Ty> d;
        return d;
    }
};

// ---- sources/ppm_probabilistic.h:23-42 -------------------------------------
class ProgressivePhotonMappingProb {
    Image _result;
    PPMSettings _settings;
    std::vector<pppm_pass_stats> _stats;
    std::function<void(int, const Image&)> _onPass;

public:
    static const double ALPHA;
    ProgressivePhotonMappingProb() {}
    explicit ProgressivePhotonMappingProb(const PPMSettings& s) : _settings(s) {}
    // Runs params.spp() passes of params.photons() photons each on the GPU.
    void render(const Scene& scene, const Camera& camera, const RenderParameters& params,
                RandomSamplerType randomSamplerType = RANDOM_SAMPLER_PSEUDO_RANDOM);
    const Image& result() const { return _result; }
    const std::vector<pppm_pass_stats>& passStats() const { return _stats; }
    // called after every pass with the running average (the reference saves a
    // PNG here, sources/ppm_probabilistic.cc:97-106)
    void setPassCallback(std::function<void(int, const Image&)> f) { _onPass = std::move(f); }
    PPMSettings& settings() { return _settings; }
};

// ---- sources/progressive_photon_mapping.h:25-108 ----------------------------
// Progressive photon mapping with a radius per render point (Hachisuka et al.
// 2008) on the same device path: pixel render points keep r^2, their photon
// count and their flux across passes.  Takes no BSSRDF materials and no
// learned proposal; settings().radiusFraction only serves a pass whose render
// points have no extent.
class ProgressivePhotonMapping {
    Image _result;
    PPMSettings _settings;
    std::vector<pppm_pass_stats> _stats;
    std::function<void(int, const Image&)> _onPass;

public:
    static const double ALPHA;
    ProgressivePhotonMapping() {}
    explicit ProgressivePhotonMapping(const PPMSettings& s) : _settings(s) {}
    void render(const Scene& scene, const Camera& camera, const RenderParameters& params);
    const Image& result() const { return _result; }
    const std::vector<pppm_pass_stats>& passStats() const { return _stats; }
    void setPassCallback(std::function<void(int, const Image&)> f) { _onPass = std::move(f); }
    PPMSettings& settings() { return _settings; }
};

}  // namespace pppm_amd
