// facade_main.cc -- drives the reference-shaped C++ interface (pppm_amd.h)
// the way the reference's application and tests drive theirs
// (sources/main/appmain.cc:11-26; tests/test_trimesh.cc:62-117).
//
//   facade_main <out_image.bin> [<sss_image.bin> <proposal_image.bin>]
// Exits 0 when every check passes; prints one line per check.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <thread>

#include "../../tatsy-pppm_amd/host/pppm_amd.h"

using namespace pppm_amd;

static int failures = 0;
#define CHECK(cond, ...)                      \
    do {                                      \
        if (!(cond)) {                        \
            std::printf("FAIL: " __VA_ARGS__); \
            std::printf("\n");                \
            failures++;                       \
        }                                     \
    } while (0)

// brute-force closest hit over the scene's triangles (Moller-Trumbore, FP64)
static int brute_force(const Scene& scene, const Ray& ray, double* tbest) {
    int hit = -1;
    *tbest = 1e20;
    for (size_t i = 0; i < scene.numTriangles(); i++) {
        const Triangle& t = scene.getTriangle((int)i);
        Vector3D e1 = t.p1() - t.p0(), e2 = t.p2() - t.p0();
        Vector3D pv = Vector3D::cross(ray.direction(), e2);
        double det = Vector3D::dot(e1, pv);
        if (std::fabs(det) < 1e-12) continue;
        Vector3D tv = ray.origin() - t.p0();
        double u = Vector3D::dot(tv, pv) / det;
        if (u < 0.0 || u > 1.0) continue;
        Vector3D qv = Vector3D::cross(tv, e1);
        double v = Vector3D::dot(ray.direction(), qv) / det;
        if (v < 0.0 || u + v > 1.0) continue;
        double tt = Vector3D::dot(e2, qv) / det;
        if (tt > 1e-12 && tt < *tbest) {
            *tbest = tt;
            hit = (int)i;
        }
    }
    return hit;
}

int main(int argc, char** argv) {
    const char* out = argc > 1 ? argv[1] : "facade_image.bin";

    // --- a scene built through Scene::add, as appmain.cc does ---------------
    Scene scene;
    PerspectiveCamera camera = scene.loadProcedural(PPPM_SCENE_BOX, 1, 0, 64, 64);
    std::printf("scene: %zu triangles\n", scene.numTriangles());
    scene.add(Triangle(Vector3D(0.5, 0.01, 0.5), Vector3D(1.0, 0.01, 0.5), Vector3D(0.5, 0.01, 1.0)),
              LambertianBRDF::factory(Vector3D(0.9, 0.1, 0.1)));
    scene.setAccelerator();

    // --- Scene::intersect against brute force (tests/test_trimesh.cc:93-116)
    std::mt19937 rng(12345);
    std::uniform_real_distribution<double> U(0.1, 4.9), N(-1.0, 1.0);
    int agree = 0, total = 0;
    for (int k = 0; k < 300; k++) {
        Vector3D o(U(rng), U(rng), U(rng));
        Vector3D d(N(rng), N(rng), N(rng));
        d = d.normalized();
        Ray ray(o, d);
        Intersection isect;
        bool h = scene.intersect(ray, isect);
        double tb;
        int b = brute_force(scene, ray, &tb);
        total++;
        if (h == (b >= 0) && (!h || std::fabs(isect.hittingDistance() - tb) < 1e-4 * (1.0 + tb))) agree++;
    }
    CHECK(agree >= total - 1, "Scene::intersect agrees with brute force on %d of %d rays", agree, total);
    std::printf("intersect: %d/%d agree\n", agree, total);

    // --- Scene::intersect from many host threads at once: const and
    // re-entrant (the reference's OpenMP loops call it concurrently,
    // ppm_probabilistic.cc:157, 256) --------------------------------------------
    {
        const int R = 4000, T = 8;
        std::vector<Ray> rays;
        for (int k = 0; k < R; k++)
            rays.push_back(Ray(Vector3D(U(rng), U(rng), U(rng)), Vector3D(N(rng), N(rng), N(rng)).normalized()));
        std::vector<int> ids(R);
        std::vector<double> ts(R);
        for (int k = 0; k < R; k++) {
            Intersection is;
            scene.intersect(rays[(size_t)k], is);
            ids[(size_t)k] = is.objectID();
            ts[(size_t)k] = is.hittingDistance();
        }
        std::vector<int> bad(T, 0);
        std::vector<std::thread> pool;
        for (int w = 0; w < T; w++)
            pool.emplace_back([&, w] {
                for (int rep = 0; rep < 3; rep++)
                    for (int k = w; k < R; k += 1 + (w % 3)) {
                        Intersection is;
                        scene.intersect(rays[(size_t)k], is);
                        if (is.objectID() != ids[(size_t)k] || is.hittingDistance() != ts[(size_t)k]) bad[(size_t)w]++;
                    }
            });
        for (std::thread& t : pool) t.join();
        int nbad = 0;
        for (int b : bad) nbad += b;
        CHECK(nbad == 0, "concurrent Scene::intersect differs from serial on %d queries", nbad);
        std::printf("concurrent intersect: %d threads, %d mismatches\n", T, nbad);
    }

    // --- QBVHAccel on its own ------------------------------------------------
    {
        std::vector<Triangle> tris;
        for (size_t i = 0; i < scene.numTriangles(); i++) tris.push_back(scene.getTriangle((int)i));
        QBVHAccel accel(scene.device());
        accel.construct(tris);
        Hitpoint hp;
        int id = accel.intersect(Ray(Vector3D(2.5, 2.5, 2.5), Vector3D(0, -1, 0)), &hp);
        CHECK(id >= 0 && std::fabs(hp.distance() - 2.5) < 1e-4, "QBVHAccel floor hit id=%d t=%f", id, hp.distance());
        Hitpoint short_ray;
        short_ray.setDistance(1.0);  // distance() is the ray's tMax on input
        int none = accel.intersect(Ray(Vector3D(2.5, 2.5, 2.5), Vector3D(0, -1, 0)), &short_ray);
        CHECK(none == -1, "QBVHAccel respects tMax");
    }

    // --- HashGrid<int> bucket contents (sources/hash_grid_detail.h) ----------
    {
        HashGrid<int> grid;
        BBox box(Vector3D(0, 0, 0), Vector3D(5, 5, 5));
        const int S = 1031;
        const double scale = 1.0 / (2.0 * 0.1);
        grid.init(S, scale, box);
        std::vector<Vector3D> pts;
        for (int i = 0; i < 2000; i++) pts.push_back(Vector3D(U(rng), U(rng), U(rng)));
        for (int i = 0; i < 2000; i++) grid.add(i, pts[i], pts[i]);
        grid.construct();
        int bad = 0;
        for (int i = 0; i < 2000; i++) {
            const std::vector<int>& b = grid[pts[i]];
            // brute force: every point in the same bucket, in insertion order
            std::vector<int> expect;
            auto bucket = [&](const Vector3D& p) {
                int c[3];
                for (int d = 0; d < 3; d++) c[d] = std::abs(static_cast<int>((p[d] - 0.0) * scale));
                return (((uint32_t)c[0] * 73856093u) ^ ((uint32_t)c[1] * 19349663u) ^ ((uint32_t)c[2] * 83492791u)) % S;
            };
            for (int j = 0; j < 2000; j++)
                if (bucket(pts[j]) == bucket(pts[i])) expect.push_back(j);
            if (b != expect) bad++;
        }
        CHECK(bad == 0, "HashGrid buckets differ for %d points", bad);
        std::printf("hashgrid: %d mismatching buckets\n", bad);
    }

    // --- the renderer (sources/main/appmain.cc:21-25) -------------------------
    {
        Scene s2;
        PerspectiveCamera cam = s2.loadProcedural(PPPM_SCENE_BOX, 1, 0, 64, 64);
        ProgressivePhotonMappingProb ppm;
        int passes_seen = 0;
        ppm.setPassCallback([&](int t, const Image&) { passes_seen = t; });
        ppm.render(s2, cam, RenderParameters(20000, 2, 0, 0.0));
        const Image& img = ppm.result();
        double sum = 0.0;
        bool finite = true;
        for (double v : img.data()) {
            sum += v;
            finite = finite && std::isfinite(v) && v >= 0.0;
        }
        CHECK(finite && sum > 0.0, "image finite and lit");
        CHECK(passes_seen == 2 && ppm.passStats().size() == 2, "two passes");
        FILE* f = std::fopen(out, "wb");
        CHECK(f != nullptr, "open %s", out);
        if (f) {
            std::fwrite(img.data().data(), sizeof(double), img.data().size(), f);
            std::fclose(f);
        }
        std::printf("render: mean %.6f, %lld photons stored in pass 2\n", sum / img.data().size(),
                    (long long)ppm.passStats()[1].photons_stored);
    }
    // --- translucent material, set up as appmain.cc:63-65 does ----------------
    if (argc > 2) {
        Scene s3;
        PerspectiveCamera cam = s3.loadProcedural(PPPM_SCENE_TWO_ROOMS_SSS, 1, 12, 48, 48);
        // the scene's translucent material rebuilt through the factories
        const BSDF& orig = s3.getBsdf((int)s3.numTriangles() - 1);
        BSDF rebuilt = RefractionBSDF::factory(orig.reflectance());
        rebuilt.setBssrdf(DipoleBSSRDF::factory(
            Vector3D(orig.raw().sigma_a[0], orig.raw().sigma_a[1], orig.raw().sigma_a[2]),
            Vector3D(orig.raw().sigmap_s[0], orig.raw().sigmap_s[1], orig.raw().sigmap_s[2]), orig.raw().eta));
        CHECK(rebuilt.type() == (BSDF_TYPE_REFRACTION | BSDF_TYPE_BSSRDF) &&
                  std::memcmp(&rebuilt.raw(), &orig.raw(), sizeof(pppm_bsdf)) == 0,
              "DipoleBSSRDF::factory + setBssrdf rebuild the scene's material");
        bool threw = false;
        try {
            DipoleBSSRDF::factory(Vector3D(0.01, 0.01, 0.01), Vector3D(1, 1, 1), 1.3, 2.0);
        } catch (const Error&) {
            threw = true;
        }
        CHECK(threw, "a dipole scale other than 1 is rejected");
        PPMSettings set;
        set.subsurfacePoints = 20000;
        ProgressivePhotonMappingProb ppm(set);
        ppm.render(s3, cam, RenderParameters(100000, 2, 0, 0.0));
        CHECK(ppm.passStats()[1].sss_hits > 0 && ppm.passStats()[1].sss_points == 20000, "subsurface estimate ran");
        FILE* f = std::fopen(argv[2], "wb");
        if (f) {
            std::fwrite(ppm.result().data().data(), sizeof(double), ppm.result().data().size(), f);
            std::fclose(f);
        }
        std::printf("subsurface: %lld eye paths in the dipole sum\n", (long long)ppm.passStats()[1].sss_hits);
    }
    // --- the learned emission proposal ---------------------------------------
    if (argc > 3) {
        Scene s4;
        PerspectiveCamera cam = s4.loadProcedural(PPPM_SCENE_TWO_ROOMS, 1, 8, 48, 48);
        PPMSettings set;
        set.learnedProposal = true;
        set.proposal.anchors = 4;
        set.proposal.train_interval = 2;
        set.proposal.train_steps = 8;
        set.proposal.batch = 4096;
        set.proposal.ring_capacity = 1 << 20;
        ProgressivePhotonMappingProb ppm(set);
        ppm.render(s4, cam, RenderParameters(50000, 4, 0, 0.0));
        CHECK(ppm.passStats()[1].retrained == 1 && ppm.passStats()[2].proposal_active == 1, "proposal trained and used");
        FILE* f = std::fopen(argv[3], "wb");
        if (f) {
            std::fwrite(ppm.result().data().data(), sizeof(double), ppm.result().data().size(), f);
            std::fclose(f);
        }
        std::printf("proposal: loss %.4f after pass 2\n", ppm.passStats()[1].train_loss);
    }
    // --- per-point radii (sources/progressive_photon_mapping.h:92-98) -----------
    if (argc > 4) {
        Scene s5;
        PerspectiveCamera cam = s5.loadProcedural(PPPM_SCENE_BOX, 1, 0, 40, 32);
        ProgressivePhotonMapping ppm;
        CHECK(ProgressivePhotonMapping::ALPHA == 0.7, "ALPHA");
        ppm.render(s5, cam, RenderParameters(30000, 3, 0, 0.0));
        CHECK(ppm.passStats().size() == 3 && ppm.passStats()[2].radius > 0.0, "three passes with a radius");
        FILE* f = std::fopen(argv[4], "wb");
        if (f) {
            std::fwrite(ppm.result().data().data(), sizeof(double), ppm.result().data().size(), f);
            std::fclose(f);
        }
        bool threw = false;
        try {
            PPMSettings set;
            set.learnedProposal = true;
            ProgressivePhotonMapping bad(set);
            bad.render(s5, cam, RenderParameters(1000, 1, 0, 0.0));
        } catch (const Error&) {
            threw = true;
        }
        CHECK(threw, "the learned proposal is refused by the per-point method");
        std::printf("hachisuka: initial radius %.6f in pass 3\n", ppm.passStats()[2].radius);
    }
    std::printf(failures ? "FAILED\n" : "OK\n");
    return failures ? 1 : 0;
}
