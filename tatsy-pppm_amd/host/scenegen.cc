// scenegen.cc -- procedural scenes for the benchmark configurations.
//
// The reference builds its scene from PLY meshes and an HDR environment map
// (sources/main/appmain.cc:29-126); the configurations here are procedural:
// a closed box, geodesic icospheres (20 f^2 triangles at frequency f) with
// optional value-noise displacement along the radius, a mirror and a glass
// sphere, two rooms joined by a slit, and a parallelogram area light.
// Everything is FP64 until the final FP32 vertex, and shared vertices are
// computed once, so meshes are watertight.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "../../include/pppm.h"

namespace {

struct D3 {
    double x, y, z;
};
D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 cross(D3 u, D3 v) { return {u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x}; }
D3 unit(D3 a) { return a * (1.0 / std::sqrt(dot(a, a))); }

// Philox-4x32-10 on the host, for seeded scene parameters.
void philox(uint32_t c[4], const uint32_t k_in[2]) {
    uint32_t k0 = k_in[0], k1 = k_in[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[0] = n0; c[1] = (uint32_t)p1; c[2] = n2; c[3] = (uint32_t)p0;
    }
}

struct SeededRng {
    uint32_t key[2];
    uint32_t stream;
    uint32_t n = 0;
    SeededRng(uint64_t seed, uint32_t s) : stream(s) {
        key[0] = (uint32_t)seed;
        key[1] = (uint32_t)(seed >> 32);
    }
    double uniform() {
        uint32_t c[4] = {n++, stream, 0x5343454Eu, 0u};
        philox(c, key);
        return (double)(c[0] >> 8) * (1.0 / 16777216.0);
    }
};

// Value noise: U(-1, 1) lattice values hashed from the seed, smoothstep
// weights, trilinear interpolation.
struct ValueNoise {
    uint32_t key[2];
    explicit ValueNoise(uint64_t seed) {
        key[0] = (uint32_t)(seed ^ 0x9E3779B97F4A7C15ull);
        key[1] = (uint32_t)((seed ^ 0x9E3779B97F4A7C15ull) >> 32);
    }
    double lattice(int64_t i, int64_t j, int64_t k) const {
        uint32_t c[4] = {(uint32_t)i, (uint32_t)j, (uint32_t)k, 0x4E4F4953u};
        philox(c, key);
        return (double)(c[0] >> 8) * (2.0 / 16777216.0) - 1.0;
    }
    double operator()(D3 p) const {
        double fx = std::floor(p.x), fy = std::floor(p.y), fz = std::floor(p.z);
        int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
        double tx = p.x - fx, ty = p.y - fy, tz = p.z - fz;
        tx = tx * tx * (3.0 - 2.0 * tx);
        ty = ty * ty * (3.0 - 2.0 * ty);
        tz = tz * tz * (3.0 - 2.0 * tz);
        double v = 0.0;
        for (int c = 0; c < 8; c++) {
            int dx = c & 1, dy = (c >> 1) & 1, dz = (c >> 2) & 1;
            double w = (dx ? tx : 1.0 - tx) * (dy ? ty : 1.0 - ty) * (dz ? tz : 1.0 - tz);
            v += w * lattice(ix + dx, iy + dy, iz + dz);
        }
        return v;
    }
};

struct Builder {
    std::vector<float> tri;
    std::vector<int32_t> mat;
    std::vector<pppm_bsdf> bsdfs;
    std::vector<pppm_light> lights;

    int add_bsdf(int type, D3 rho, D3 le = {0, 0, 0}) {
        pppm_bsdf b;
        std::memset(&b, 0, sizeof(b));
        b.type = type;
        b.reflectance[0] = (float)rho.x; b.reflectance[1] = (float)rho.y; b.reflectance[2] = (float)rho.z;
        b.emission[0] = (float)le.x; b.emission[1] = (float)le.y; b.emission[2] = (float)le.z;
        b.ior = 1.5f;
        b.eta = 1.3f;
        bsdfs.push_back(b);
        return (int)bsdfs.size() - 1;
    }
    void add_tri(D3 a, D3 b, D3 c, int m) {
        const D3 v[3] = {a, b, c};
        for (int k = 0; k < 3; k++) {
            tri.push_back((float)v[k].x);
            tri.push_back((float)v[k].y);
            tri.push_back((float)v[k].z);
        }
        mat.push_back(m);
    }
    // rectangle corner + eu + ev, two triangles with normal eu x ev
    void add_quad(D3 c, D3 eu, D3 ev, int m) {
        add_tri(c, c + eu, c + eu + ev, m);
        add_tri(c, c + eu + ev, c + ev, m);
    }
    // axis-aligned box walls, one material per wall (-x, +x, -y, +y, -z, +z);
    // winding is irrelevant: only emitters have a front face
    void add_box(D3 lo, D3 hi, const int m[6], const bool skip[6] = nullptr) {
        D3 ex = {hi.x - lo.x, 0, 0}, ey = {0, hi.y - lo.y, 0}, ez = {0, 0, hi.z - lo.z};
        if (!skip || !skip[0]) add_quad(lo, ez, ey, m[0]);
        if (!skip || !skip[1]) add_quad(lo + ex, ey, ez, m[1]);
        if (!skip || !skip[2]) add_quad(lo, ex, ez, m[2]);
        if (!skip || !skip[3]) add_quad(lo + ey, ez, ex, m[3]);
        if (!skip || !skip[4]) add_quad(lo, ey, ex, m[4]);
        if (!skip || !skip[5]) add_quad(lo + ez, ex, ey, m[5]);
    }
    // downward-facing (-y) square light of side `size` centred at (cx, y, cz)
    void add_ceiling_light(double cx, double y, double cz, double size, D3 le) {
        D3 c = {cx - 0.5 * size, y, cz - 0.5 * size};
        D3 eu = {size, 0, 0}, ev = {0, 0, size};  // eu x ev = (0, -size^2, 0)
        int m = add_bsdf(PPPM_BSDF_EMITTER, {0, 0, 0}, le);
        add_quad(c, eu, ev, m);
        pppm_light L;
        L.corner[0] = (float)c.x; L.corner[1] = (float)c.y; L.corner[2] = (float)c.z;
        L.edge_u[0] = (float)eu.x; L.edge_u[1] = (float)eu.y; L.edge_u[2] = (float)eu.z;
        L.edge_v[0] = (float)ev.x; L.edge_v[1] = (float)ev.y; L.edge_v[2] = (float)ev.z;
        L.emission[0] = (float)le.x; L.emission[1] = (float)le.y; L.emission[2] = (float)le.z;
        lights.push_back(L);
    }
    // geodesic icosphere of frequency f: 20 f^2 triangles
    void add_icosphere(D3 center, double radius, int f, int m, const ValueNoise* noise,
                       double disp_amp, double disp_freq) {
        const double t = (1.0 + std::sqrt(5.0)) / 2.0;
        const D3 iv[12] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                           {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
        int faces[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11},
                            {1, 5, 9}, {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                            {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8}, {3, 8, 9},
                            {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
        D3 V[12];
        for (int i = 0; i < 12; i++) V[i] = unit(iv[i]);
        // canonical key: (vertex id, weight) pairs with nonzero weight, sorted by id
        typedef std::tuple<int, int, int, int, int, int> Key;
        std::map<Key, int> index;
        std::vector<D3> verts;
        auto vertex = [&](int a, int wa, int b, int wb, int c, int wc) -> int {
            std::pair<int, int> p[3] = {{a, wa}, {b, wb}, {c, wc}};
            std::pair<int, int> q[3];
            int nq = 0;
            for (auto& e : p)
                if (e.second != 0) q[nq++] = e;
            for (int i = 1; i < nq; i++)
                for (int j = i; j > 0 && q[j].first < q[j - 1].first; j--) std::swap(q[j], q[j - 1]);
            Key k{nq > 0 ? q[0].first : -1, nq > 0 ? q[0].second : 0, nq > 1 ? q[1].first : -1,
                  nq > 1 ? q[1].second : 0, nq > 2 ? q[2].first : -1, nq > 2 ? q[2].second : 0};
            auto it = index.find(k);
            if (it != index.end()) return it->second;
            D3 s = {0, 0, 0};
            for (int i = 0; i < nq; i++) s = s + V[q[i].first] * (double)q[i].second;
            D3 u = unit(s * (1.0 / f));
            double r = radius;
            if (noise) r += disp_amp * (*noise)(u * (disp_freq * radius) + center);
            verts.push_back(center + u * r);
            int id = (int)verts.size() - 1;
            index.emplace(k, id);
            return id;
        };
        for (auto& fc : faces) {
            int A = fc[0], B = fc[1], C = fc[2];
            if (dot(cross(V[B] - V[A], V[C] - V[A]), V[A] + V[B] + V[C]) < 0.0) std::swap(B, C);
            for (int i = 0; i < f; i++)
                for (int j = 0; j < f - i; j++) {
                    // P(i, j) = ((f - i - j) A + i B + j C) / f
                    int p00 = vertex(A, f - i - j, B, i, C, j);
                    int p10 = vertex(A, f - i - 1 - j, B, i + 1, C, j);
                    int p01 = vertex(A, f - i - j - 1, B, i, C, j + 1);
                    add_tri(verts[p00], verts[p10], verts[p01], m);
                    if (j < f - 1 - i) {
                        int p11 = vertex(A, f - i - j - 2, B, i + 1, C, j + 1);
                        add_tri(verts[p10], verts[p11], verts[p01], m);
                    }
                }
        }
    }
};

D3 albedo(SeededRng& r) {
    double a = 0.2 + 0.6 * r.uniform();
    double b = 0.2 + 0.6 * r.uniform();
    double c = 0.2 + 0.6 * r.uniform();
    return {a, b, c};
}

void finish(Builder& b, pppm_scene_desc* out, pppm_camera cam) {
    out->n_tris = (int64_t)b.mat.size();
    out->tri_xyz = (float*)std::malloc(sizeof(float) * b.tri.size() + 16);
    std::memcpy(out->tri_xyz, b.tri.data(), sizeof(float) * b.tri.size());
    out->bsdf_id = (int32_t*)std::malloc(sizeof(int32_t) * b.mat.size() + 16);
    std::memcpy(out->bsdf_id, b.mat.data(), sizeof(int32_t) * b.mat.size());
    out->n_bsdf = (int32_t)b.bsdfs.size();
    out->bsdfs = (pppm_bsdf*)std::malloc(sizeof(pppm_bsdf) * b.bsdfs.size() + 16);
    std::memcpy(out->bsdfs, b.bsdfs.data(), sizeof(pppm_bsdf) * b.bsdfs.size());
    out->n_lights = (int32_t)b.lights.size();
    out->lights = (pppm_light*)std::malloc(sizeof(pppm_light) * b.lights.size() + 16);
    std::memcpy(out->lights, b.lights.data(), sizeof(pppm_light) * b.lights.size());
    out->camera = cam;
    for (int d = 0; d < 3; d++) {
        out->bbox_min[d] = 1e300;
        out->bbox_max[d] = -1e300;
    }
    for (size_t i = 0; i < b.tri.size(); i++) {
        out->bbox_min[i % 3] = std::fmin(out->bbox_min[i % 3], (double)b.tri[i]);
        out->bbox_max[i % 3] = std::fmax(out->bbox_max[i % 3], (double)b.tri[i]);
    }
}

pppm_camera make_cam(D3 c, D3 d, double fov, int w, int h) {
    pppm_camera cam;
    cam.center[0] = c.x; cam.center[1] = c.y; cam.center[2] = c.z;
    cam.direction[0] = d.x; cam.direction[1] = d.y; cam.direction[2] = d.z;
    cam.up[0] = 0.0; cam.up[1] = 1.0; cam.up[2] = 0.0;
    cam.fov_y_deg = fov;
    cam.width = w;
    cam.height = h;
    return cam;
}

}  // namespace

extern "C" {

int pppm_scenegen(int32_t kind, uint64_t seed, int32_t frequency, int32_t width, int32_t height,
                  pppm_scene_desc* out) {
    if (!out || width <= 0 || height <= 0 || frequency < 0 || frequency > 2000) return PPPM_ERR_INVALID;
    std::memset(out, 0, sizeof(*out));
    Builder b;
    SeededRng rng(seed, (uint32_t)kind);
    ValueNoise noise(seed);
    const D3 le = {10.0, 10.0, 10.0};
    if (kind == PPPM_SCENE_BOX || kind == PPPM_SCENE_BOX_CAUSTIC || kind == PPPM_SCENE_RANDOM) {
        // closed box [0,5]^3, one Lambertian albedo U(0.2, 0.8)^3 per wall
        int wm[6];
        for (int k = 0; k < 6; k++) wm[k] = b.add_bsdf(PPPM_BSDF_LAMBERTIAN, albedo(rng));
        b.add_box({0, 0, 0}, {5, 5, 5}, wm);
        b.add_ceiling_light(2.5, 5.0 - 1e-3, 2.5, 1.0, le);
        if (kind == PPPM_SCENE_BOX) {
            int f = frequency > 0 ? frequency : 8;  // 1,280 triangles per sphere
            int s0 = b.add_bsdf(PPPM_BSDF_LAMBERTIAN, albedo(rng));
            int s1 = b.add_bsdf(PPPM_BSDF_LAMBERTIAN, albedo(rng));
            b.add_icosphere({1.5, 1.0, 1.8}, 1.0, f, s0, nullptr, 0.0, 0.0);
            b.add_icosphere({3.5, 1.0, 3.0}, 1.0, f, s1, nullptr, 0.0, 0.0);
        } else if (kind == PPPM_SCENE_BOX_CAUSTIC) {
            int f = frequency > 0 ? frequency : 128;  // 327,680 triangles per sphere
            int s0 = b.add_bsdf(PPPM_BSDF_LAMBERTIAN, albedo(rng));
            int s1 = b.add_bsdf(PPPM_BSDF_LAMBERTIAN, albedo(rng));
            int mir = b.add_bsdf(PPPM_BSDF_SPECULAR, {0.999, 0.999, 0.999});
            int gls = b.add_bsdf(PPPM_BSDF_REFRACTION, {0.999, 0.999, 0.999});
            b.add_icosphere({1.3, 1.0, 1.4}, 1.0, f, s0, &noise, 0.05, 4.0);
            b.add_icosphere({3.7, 1.0, 1.4}, 1.0, f, s1, &noise, 0.05, 4.0);
            int fs = f >= 32 ? 32 : f;  // 20,480 triangles each
            b.add_icosphere({1.4, 0.7, 3.3}, 0.7, fs, mir, nullptr, 0.0, 0.0);
            b.add_icosphere({3.6, 0.7, 3.3}, 0.7, fs, gls, nullptr, 0.0, 0.0);
        } else {
            // one scene of a batch: a single sphere of the requested frequency
            int f = frequency > 0 ? frequency : 8;
            int s0 = b.add_bsdf(PPPM_BSDF_LAMBERTIAN, albedo(rng));
            double cx = 1.5 + 2.0 * rng.uniform(), cz = 1.5 + 2.0 * rng.uniform();
            b.add_icosphere({cx, 1.2, cz}, 1.2, f, s0, &noise, 0.05, 4.0);
        }
        finish(b, out, make_cam({2.5, 2.5, 4.9}, {0.0, -0.12, -1.0}, 60.0, width, height));
        return PPPM_OK;
    }
    if (kind == PPPM_SCENE_TWO_ROOMS || kind == PPPM_SCENE_TWO_ROOMS_SSS) {
        // room A: x in [0, 6], room B: x in [6, 12]; shared wall x = 6 with a
        // slit z in [2.9, 3.1], y in [0, 2]
        int wa[6], wb[6];
        for (int k = 0; k < 6; k++) wa[k] = b.add_bsdf(PPPM_BSDF_LAMBERTIAN, albedo(rng));
        for (int k = 0; k < 6; k++) wb[k] = b.add_bsdf(PPPM_BSDF_LAMBERTIAN, albedo(rng));
        bool skip_a[6] = {false, true, false, false, false, false};
        bool skip_b[6] = {true, false, false, false, false, false};
        b.add_box({0, 0, 0}, {6, 6, 6}, wa, skip_a);
        b.add_box({6, 0, 0}, {12, 6, 6}, wb, skip_b);
        int wall = b.add_bsdf(PPPM_BSDF_LAMBERTIAN, albedo(rng));
        b.add_quad({6, 0, 0}, {0, 6, 0}, {0, 0, 2.9}, wall);
        b.add_quad({6, 0, 3.1}, {0, 6, 0}, {0, 0, 2.9}, wall);
        b.add_quad({6, 2, 2.9}, {0, 4, 0}, {0, 0, 0.2}, wall);
        b.add_ceiling_light(3.0, 6.0 - 1e-3, 3.0, 1.0, le);
        int f = frequency > 0 ? frequency : (kind == PPPM_SCENE_TWO_ROOMS ? 100 : 316);
        int sm;
        if (kind == PPPM_SCENE_TWO_ROOMS_SSS) {
            // translucent: specular dielectric boundary over a dipole BSSRDF
            // (the marble set-up of sources/main/appmain.cc:43-66 with the
            // coefficients drawn per channel)
            sm = b.add_bsdf(PPPM_BSDF_REFRACTION | PPPM_BSDF_BSSRDF, {0.99, 0.99, 0.99});
            for (int k = 0; k < 3; k++) {
                b.bsdfs[(size_t)sm].sigma_a[k] = (float)(0.001 + 0.049 * rng.uniform());
                b.bsdfs[(size_t)sm].sigmap_s[k] = (float)(0.5 + 1.5 * rng.uniform());
            }
        } else {
            sm = b.add_bsdf(PPPM_BSDF_LAMBERTIAN, albedo(rng));
        }
        b.add_icosphere({9.0, 1.3, 3.0}, 1.3, f, sm, &noise, 0.05, 4.0);
        finish(b, out, make_cam({11.9, 3.0, 3.0}, {-1.0, -0.15, 0.0}, 60.0, width, height));
        return PPPM_OK;
    }
    return PPPM_ERR_INVALID;
}

void pppm_scenegen_free(pppm_scene_desc* d) {
    if (!d) return;
    std::free(d->tri_xyz);
    std::free(d->bsdf_id);
    std::free(d->bsdfs);
    std::free(d->lights);
    std::memset(d, 0, sizeof(*d));
}

}  // extern "C"
