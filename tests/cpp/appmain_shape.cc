// appmain_shape.cc -- the reference application's call pattern compiled
// against the C++ interface (sources/main/appmain.cc:11-26): a default
// Camera assigned inside setScene, a mesh read from a .ply file and placed
// with fitToBBox / scale / putOnPlane / translate, RenderParameters(photons,
// spp, gatherPhotons, gatherRadius), and render(scene, camera, params) with
// the default random sampler.  Only the emitter differs: an area light
// (addAreaLight) where the reference loads an environment map.
//
//   appmain_shape [width] [height] [spp] [mesh.ply]
// Writes the mesh file it loads (an icosphere) when none is given; exits 0 and
// prints OK when the render is finite, non-negative and lit.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../tatsy-pppm_amd/host/pppm_amd.h"

using namespace pppm_amd;

static std::string g_mesh = "appmain_shape_mesh.ply";

// An icosphere written as binary little-endian PLY (float x y z, uchar + int
// face lists): the layout of the reference's own assets.
static void write_sphere_ply(const std::string& path, int subdiv) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<Vector3D> v = {{-1, t, 0}, {1, t, 0},  {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                               {0, -1, -t}, {0, 1, -t}, {t, 0, -1},  {t, 0, 1},   {-t, 0, -1}, {-t, 0, 1}};
    std::vector<Triplet> f = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                              {11, 10, 2}, {10, 7, 6}, {7, 1, 8},  {3, 9, 4},  {3, 4, 2},   {3, 2, 6}, {3, 6, 8},
                              {3, 8, 9},  {4, 9, 5},  {2, 4, 11}, {6, 2, 10}, {8, 6, 7},   {9, 8, 1}};
    for (Vector3D& p : v) p = p.normalized();
    for (int s = 0; s < subdiv; s++) {
        std::vector<Triplet> g;
        for (const Triplet& q : f) {
            int m[3];
            for (int e = 0; e < 3; e++) {
                v.push_back(((v[(size_t)q[e]] + v[(size_t)q[(e + 1) % 3]]) * 0.5).normalized());
                m[e] = (int)v.size() - 1;
            }
            g.push_back(Triplet(q[0], m[0], m[2]));
            g.push_back(Triplet(q[1], m[1], m[0]));
            g.push_back(Triplet(q[2], m[2], m[1]));
            g.push_back(Triplet(m[0], m[1], m[2]));
        }
        f = g;
    }
    FILE* fp = std::fopen(path.c_str(), "wb");
    if (!fp) {
        std::printf("cannot write %s\n", path.c_str());
        std::exit(2);
    }
    std::fprintf(fp, "ply\nformat binary_little_endian 1.0\nelement vertex %zu\nproperty float x\nproperty float y\n"
                     "property float z\nelement face %zu\nproperty list uchar int vertex_indices\nend_header\n",
                 v.size(), f.size());
    for (const Vector3D& p : v) {
        const float xyz[3] = {(float)p.x(), (float)p.y(), (float)p.z()};
        std::fwrite(xyz, 4, 3, fp);
    }
    for (const Triplet& q : f) {
        const unsigned char n = 3;
        const int idx[3] = {q[0], q[1], q[2]};
        std::fwrite(&n, 1, 1, fp);
        std::fwrite(idx, 4, 3, fp);
    }
    std::fclose(fp);
}

//! Hard-coded render scene
void setScene(Scene* scene, Camera* camera, int imageWidth, int imageHeight);

static int check(const ProgressivePhotonMappingProb& ppm, const Camera& camera);

//! Main
int main(int argc, char** argv) {
    const int imageWidth  = argc >= 2 ? atoi(argv[1]) : 64;
    const int imageHeight = argc >= 3 ? atoi(argv[2]) : 48;
    const int spp         = argc >= 4 ? atoi(argv[3]) : 2;
    if (argc >= 5) g_mesh = argv[4];
    else write_sphere_ply(g_mesh, 3);

    Scene scene;
    Camera camera;
    setScene(&scene, &camera, imageWidth, imageHeight);

    // Set render parameters
    RenderParameters params(20000, spp, 128, 16.0);

    // Set renderer
    ProgressivePhotonMappingProb ppmapa;
    ppmapa.render(scene, camera, params);
    return check(ppmapa, camera);
}

//! Hard-coded render scene: a diffuse room, a placed mesh, an area light
void setScene(Scene* scene, Camera* camera, int imageWidth, int imageHeight) {
    std::printf("Preparing the scene -> ");

    Trimesh trimesh;
    trimesh.load(g_mesh);
    trimesh.fitToBBox(BBox(Vector3D(-10.0, -10.0, -10.0), Vector3D(10.0, 10.0, 10.0)));
    trimesh.scale(0.2);
    trimesh.putOnPlane(Plane(10.0, Vector3D(0.0, 1.0, 0.0)));
    trimesh.translate(Vector3D(1.0, 0.0, -1.0));
    scene->add(trimesh, LambertianBRDF::factory(Vector3D(0.75, 0.75, 0.75)));

    // a room [-10, 10]^3 around the mesh, floor at y = -10
    const double L = 10.0;
    const Vector3D c[8] = {{-L, -L, -L}, {L, -L, -L}, {L, L, -L}, {-L, L, -L},
                           {-L, -L, L},  {L, -L, L},  {L, L, L},  {-L, L, L}};
    const int quads[6][4] = {{0, 1, 2, 3}, {4, 7, 6, 5}, {0, 4, 5, 1}, {3, 2, 6, 7}, {0, 3, 7, 4}, {1, 5, 6, 2}};
    const Vector3D colors[6] = {{0.7, 0.7, 0.7}, {0.7, 0.7, 0.7}, {0.8, 0.8, 0.8},
                                {0.8, 0.8, 0.8}, {0.7, 0.3, 0.3}, {0.3, 0.3, 0.7}};
    for (int q = 0; q < 6; q++) {
        BSDF bsdf = LambertianBRDF::factory(colors[q]);
        scene->add(Triangle(c[quads[q][0]], c[quads[q][1]], c[quads[q][2]]), bsdf);
        scene->add(Triangle(c[quads[q][0]], c[quads[q][2]], c[quads[q][3]]), bsdf);
    }
    scene->addAreaLight(Vector3D(-2.0, L - 0.01, -2.0), Vector3D(4.0, 0.0, 0.0), Vector3D(0.0, 0.0, 4.0),
                        Vector3D(20.0, 20.0, 20.0));

    // Set accelerator
    scene->setAccelerator();

    // Set camera
    Vector3D eye(0.0, 0.0, 9.5);
    *camera = Camera(eye, -eye.normalized(), Vector3D(0.0, 1.0, 0.0), 60.0, imageWidth, imageHeight, 1.0);

    std::printf("OK\n");
}

static int check(const ProgressivePhotonMappingProb& ppm, const Camera& camera) {
    const Image& img = ppm.result();
    int fails = 0;
    auto expect = [&](bool cond, const char* what) {
        std::printf("%s: %s\n", what, cond ? "ok" : "FAILED");
        fails += cond ? 0 : 1;
    };
    expect(img.width() == camera.imagesize().width() && img.height() == camera.imagesize().height(),
           "image size is the camera's");
    double sum = 0.0, lo = 0.0;
    bool finite = true;
    for (double v : img.data()) {
        finite = finite && std::isfinite(v);
        lo = std::min(lo, v);
        sum += v;
    }
    expect(finite, "image finite");
    // compensated sums of non-negative terms: at worst a rounding below zero
    expect(lo > -1e-6 * (sum / img.data().size()), "image non-negative");
    expect(sum > 0.0 && !ppm.passStats().empty(), "image lit");
    // the camera ray through the image centre is the view direction
    const Ray r = camera.getRay(0.5 * camera.imagesize().width(), 0.5 * camera.imagesize().height());
    expect(std::fabs(Vector3D::dot(r.direction(), camera.direction().normalized()) - 1.0) < 1e-9,
           "centre ray along the view direction");
    // a quasi-Monte Carlo sampler is refused with an error, not a silent substitute
    bool refused = false;
    try {
        ProgressivePhotonMappingProb other;
        Scene empty;
        other.render(empty, camera, RenderParameters(100, 1, 16, 16.0), RANDOM_SAMPLER_QUASI_MONTE_CARLO);
    } catch (const Error& e) {
        refused = e.code == PPPM_ERR_INVALID;
    }
    expect(refused, "quasi-Monte Carlo sampler refused");
    std::printf("mean %.6g (min %.3g) over %dx%d: %s\n", sum / img.data().size(), lo, img.width(), img.height(),
                fails ? "FAIL" : "OK");
    return fails ? 1 : 0;
}
