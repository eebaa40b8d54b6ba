// ref_driver.cc -- C entry points over the compiled tatsy-pppm reference.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile (target `ref`) compiles this file
// together with the reference's own sources from /root/reference/sources into
// oracle/_ref/libref.so.  Tests use it to pin the oracle's restatements
// against the real reference functions; nothing in the product loads it.
#include <cstdint>
#include <vector>

#include "bbox.h"
#include "brdf.h"
#include "bsdf.h"
#include "bssrdf.h"
#include "hash_grid.h"
#include "perspective_camera.h"
#include "qbvh_accel.h"
#include "ray.h"
#include "reflectance.h"
#include "sampler.h"
#include "triangle.h"
#include "trimesh.h"
#include "image.h"
#include "envmap.h"
#include "random_sequence.h"
#include "plane.h"

// SubsurfaceIntegrator's octree (subsurface_integrator.cc:69-167) is a private
// member type: the driver reads that one header with private members opened
// (every standard header it pulls in is included above first).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#define private public
#include "subsurface_integrator.h"
#undef private

void* enabler = nullptr;  // declared extern in common.h:41

extern "C" {

// HashGrid<int> (hash_grid_detail.h): add(i, p_i, p_i) for every point, then
// for each point write its bucket (operator[]) as a list into `lists`, laid
// out back to back; list_off[i] .. list_off[i+1].  Returns total length, or -1
// if it would exceed cap.
int64_t ref_hashgrid_point_buckets(const float* pos, int64_t n, int32_t hash_size,
                                   double hash_scale, const double bbox_min[3],
                                   const double bbox_max[3], int32_t* lists, int64_t cap,
                                   int64_t* list_off) {
    HashGrid<int> grid;
    BBox box(Vector3D(bbox_min[0], bbox_min[1], bbox_min[2]),
             Vector3D(bbox_max[0], bbox_max[1], bbox_max[2]));
    grid.init(hash_size, hash_scale, box);
    for (int64_t i = 0; i < n; i++) {
        Vector3D p(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        grid.add((int)i, p, p);
    }
    int64_t k = 0;
    list_off[0] = 0;
    for (int64_t i = 0; i < n; i++) {
        Vector3D p(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        const std::vector<int>& b = grid[p];
        if (k + (int64_t)b.size() > cap) return -1;
        for (size_t j = 0; j < b.size(); j++) lists[k++] = b[j];
        list_off[i + 1] = k;
    }
    return k;
}

// Same with box insertion add(i, bmin_i, bmax_i); query points q_j return
// their bucket lists.
int64_t ref_hashgrid_box_buckets(const double* box_min, const double* box_max, int64_t n,
                                 int32_t hash_size, double hash_scale, const double bbox_min[3],
                                 const double bbox_max[3], const double* query, int64_t nq,
                                 int32_t* lists, int64_t cap, int64_t* list_off) {
    HashGrid<int> grid;
    BBox box(Vector3D(bbox_min[0], bbox_min[1], bbox_min[2]),
             Vector3D(bbox_max[0], bbox_max[1], bbox_max[2]));
    grid.init(hash_size, hash_scale, box);
    for (int64_t i = 0; i < n; i++) {
        grid.add((int)i, Vector3D(box_min[3 * i], box_min[3 * i + 1], box_min[3 * i + 2]),
                 Vector3D(box_max[3 * i], box_max[3 * i + 1], box_max[3 * i + 2]));
    }
    int64_t k = 0;
    list_off[0] = 0;
    for (int64_t j = 0; j < nq; j++) {
        const std::vector<int>& b = grid[Vector3D(query[3 * j], query[3 * j + 1], query[3 * j + 2])];
        if (k + (int64_t)b.size() > cap) return -1;
        for (size_t m = 0; m < b.size(); m++) lists[k++] = b[m];
        list_off[j + 1] = k;
    }
    return k;
}

// QBVHAccel::construct + intersect (qbvh_accel.cc:139-285) over FP32
// triangles; rays [n][8] as pppm_ray (tmin/tmax ignored: the reference has
// tMin = 0 and a fresh Hitpoint per query).
void ref_qbvh_intersect(const float* tri, int64_t ntri, const float* rays, int64_t nrays,
                        double* t_out, int32_t* id_out) {
    std::vector<Triangle> tris;
    tris.reserve(ntri);
    for (int64_t i = 0; i < ntri; i++) {
        const float* t = tri + 9 * i;
        tris.push_back(Triangle(Vector3D(t[0], t[1], t[2]), Vector3D(t[3], t[4], t[5]),
                                Vector3D(t[6], t[7], t[8])));
    }
    QBVHAccel accel;
    accel.construct(tris);
    for (int64_t i = 0; i < nrays; i++) {
        const float* r = rays + 8 * i;
        Vector3D d(r[4], r[5], r[6]);
        Ray ray(Vector3D(r[0], r[1], r[2]), d);
        Hitpoint hp;
        int id = accel.intersect(ray, &hp);
        id_out[i] = id;
        t_out[i] = hp.distance();
    }
}

// Triangle::intersect (triangle.cc:53-78).  Returns 1 on hit.
int32_t ref_triangle_intersect(const double tri[9], const double o[3], const double d[3],
                               double* t, double nrm[3]) {
    Triangle T(Vector3D(tri[0], tri[1], tri[2]), Vector3D(tri[3], tri[4], tri[5]),
               Vector3D(tri[6], tri[7], tri[8]));
    Ray ray(Vector3D(o[0], o[1], o[2]), Vector3D(d[0], d[1], d[2]));
    Hitpoint hp;
    bool h = T.intersect(ray, &hp);
    *t = hp.distance();
    nrm[0] = hp.normal().x(); nrm[1] = hp.normal().y(); nrm[2] = hp.normal().z();
    return h ? 1 : 0;
}

// sampler::onHemisphere (sampler.cc:17-32)
void ref_on_hemisphere(const double n[3], double r1, double r2, double out[3]) {
    Vector3D d;
    sampler::onHemisphere(Vector3D(n[0], n[1], n[2]), &d, r1, r2);
    out[0] = d.x(); out[1] = d.y(); out[2] = d.z();
}

// checkTotalReflection (reflectance.h:7-34)
int32_t ref_total_reflection(int32_t into, const double in[3], const double n[3],
                             const double on[3], double refl[3], double refr[3], double* fre,
                             double* ftr) {
    Vector3D rd, td;
    bool tr = checkTotalReflection(into != 0, Vector3D(in[0], in[1], in[2]),
                                   Vector3D(n[0], n[1], n[2]), Vector3D(on[0], on[1], on[2]),
                                   &rd, &td, fre, ftr);
    refl[0] = rd.x(); refl[1] = rd.y(); refl[2] = rd.z();
    refr[0] = td.x(); refr[1] = td.y(); refr[2] = td.z();
    return tr ? 1 : 0;
}

// BSDF::sample (bsdf.cc:62-64 -> brdf.cc).  type: BsdfType bit of the BSDF.
void ref_bsdf_sample(int32_t type, const double rho[3], const double in[3], const double n[3],
                     double r1, double r2, double out[3], double* pdf) {
    Vector3D refl(rho[0], rho[1], rho[2]);
    BSDF b;
    if (type == BSDF_TYPE_LAMBERTIAN_BRDF) b = LambertianBRDF::factory(refl);
    else if (type == BSDF_TYPE_SPECULAR_BRDF) b = SpecularBRDF::factory(refl);
    else b = RefractionBSDF::factory(refl);
    Vector3D o;
    b.sample(Vector3D(in[0], in[1], in[2]), Vector3D(n[0], n[1], n[2]), r1, r2, &o, pdf);
    out[0] = o.x(); out[1] = o.y(); out[2] = o.z();
}

// PerspectiveCamera::getRay (perspective_camera.cc:42-46); fov in the unit
// the reference uses (radians).
void ref_camera_ray(const double c[3], const double d[3], const double up[3], double fov,
                    int32_t w, int32_t h, double px, double py, double o_out[3],
                    double d_out[3]) {
    PerspectiveCamera cam(Vector3D(c[0], c[1], c[2]), Vector3D(d[0], d[1], d[2]),
                          Vector3D(up[0], up[1], up[2]), fov, w, h, 1.0);
    Ray r = cam.getRay(px, py);
    o_out[0] = r.origin().x(); o_out[1] = r.origin().y(); o_out[2] = r.origin().z();
    d_out[0] = r.direction().x(); d_out[1] = r.direction().y(); d_out[2] = r.direction().z();
}

// DipoleBSSRDF::operator() (bssrdf.cc:92-104) and Fdr (bssrdf.cc:33-42)
void ref_dipole(const double sigma_a[3], const double sigmap_s[3], double eta, double d2,
                double out[3], double* fdr) {
    BSSRDF b = DipoleBSSRDF::factory(Vector3D(sigma_a[0], sigma_a[1], sigma_a[2]),
                                     Vector3D(sigmap_s[0], sigmap_s[1], sigmap_s[2]), eta);
    Vector3D v = b(d2);
    out[0] = v.x(); out[1] = v.y(); out[2] = v.z();
    *fdr = b.Fdr();
}

// Trimesh::load + the placement of sources/main/appmain.cc:34-38 (fitToBBox,
// scale, putOnPlane, translate); writes the triangles [n][9] in FP64 and
// returns their count (-1 above cap).  do_place = 0: the triangles as loaded.
int64_t ref_trimesh_place(const char* path, int32_t do_place, const double box[6], double scale_all,
                          double plane_d, const double plane_n[3], const double move[3], double* tris,
                          int64_t cap) {
    Trimesh m;
    m.load(path);
    if (do_place) {
        m.fitToBBox(BBox(box[0], box[1], box[2], box[3], box[4], box[5]));
        m.scale(scale_all);
        m.putOnPlane(Plane(plane_d, Vector3D(plane_n[0], plane_n[1], plane_n[2])));
        m.translate(Vector3D(move[0], move[1], move[2]));
    }
    std::vector<Triangle> t = m.triangulate();
    if ((int64_t)t.size() > cap) return -1;
    for (size_t i = 0; i < t.size(); i++)
        for (int k = 0; k < 3; k++)
            for (int d = 0; d < 3; d++) tris[9 * i + 3 * k + d] = t[i].p(k)[d];
    return (int64_t)t.size();
}

// Image::save of an FP64 [h][w][3] buffer (top row first) by extension
// (.png / .bmp through the reference's vendored writer, .hdr its own RGBE).
void ref_image_save(const char* path, const double* rgb, int32_t w, int32_t h) {
    Image img(w, h);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const double* p = rgb + 3 * ((size_t)y * w + x);
            img.pixel(x, y) = Vector3D(p[0], p[1], p[2]);
        }
    img.save(path);
}

// Image::load (.hdr or .bmp) into out [h][w][3]; returns w * h, -1 above cap.
int64_t ref_image_load(const char* path, double* out, int64_t cap, int32_t* w, int32_t* h) {
    Image img;
    img.load(path);
    *w = (int32_t)img.width();
    *h = (int32_t)img.height();
    const int64_t n = (int64_t)img.width() * img.height();
    if (n > cap) return -1;
    for (int y = 0; y < (int)img.height(); y++)
        for (int x = 0; x < (int)img.width(); x++) {
            const Vector3D& c = img(x, y);
            double* p = out + 3 * ((size_t)y * img.width() + x);
            p[0] = c.x();
            p[1] = c.y();
            p[2] = c.z();
        }
    return n;
}

// Envmap(path): samplePhoton with the given uniforms (r[3 k .. 3 k + 2]: CDF
// pick, u jitter, v jitter) for k < n, photon count N: position, flux,
// direction; and sampleFromDir of the given directions.
void ref_envmap_sample(const char* path, const double* r, int64_t n, int32_t N, double* pos, double* flux,
                       double* dir) {
    Envmap env(path);
    for (int64_t k = 0; k < n; k++) {
        RandomSequence rs;
        rs.resize(3);
        for (int j = 0; j < 3; j++) rs.set(j, r[3 * k + j]);
        const Photon ph = env.samplePhoton(rs, N);
        for (int d = 0; d < 3; d++) {
            pos[3 * k + d] = ph[d];
            flux[3 * k + d] = ph.flux()[d];
            dir[3 * k + d] = ph.direction()[d];
        }
    }
}

void ref_envmap_lookup(const char* path, const double* dirs, int64_t n, double* rgb) {
    Envmap env(path);
    for (int64_t k = 0; k < n; k++) {
        const Vector3D c = env.sampleFromDir(Vector3D(dirs[3 * k], dirs[3 * k + 1], dirs[3 * k + 2]));
        rgb[3 * k] = c.x();
        rgb[3 * k + 1] = c.y();
        rgb[3 * k + 2] = c.z();
    }
}

// SubsurfaceIntegrator::Octree::construct + iradSubsurface with the reference's
// maximum error, and irradiance()'s factor (subsurface_integrator.cc:234-238):
// out[i] = (1 - F_dr) / pi * M_o(q_i) for points of equal area
void ref_dipole_octree(const double sigma_a[3], const double sigmap_s[3], double eta, const float* pts,
                       const float* E, int64_t n, double area, const float* q, int64_t nq, double max_error,
                       double* out) {
    SubsurfaceIntegrator integ;
    integ._maxError = max_error;
    std::vector<IrradiancePoint> ip((size_t)n);
    for (int64_t i = 0; i < n; i++)
        ip[(size_t)i] = IrradiancePoint(Vector3D(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]), Vector3D(0.0, 0.0, 1.0),
                                        area, Vector3D(E[3 * i], E[3 * i + 1], E[3 * i + 2]));
    SubsurfaceIntegrator::Octree oct;
    oct.construct(&integ, ip);
    BSSRDF b = DipoleBSSRDF::factory(Vector3D(sigma_a[0], sigma_a[1], sigma_a[2]),
                                     Vector3D(sigmap_s[0], sigmap_s[1], sigmap_s[2]), eta);
    for (int64_t i = 0; i < nq; i++) {
        const Vector3D Mo = oct.iradSubsurface(Vector3D(q[3 * i], q[3 * i + 1], q[3 * i + 2]), b);
        const Vector3D L = (1.0 / PI) * (1.0 - b.Fdr()) * Mo;
        out[3 * i] = L.x();
        out[3 * i + 1] = L.y();
        out[3 * i + 2] = L.z();
    }
}

}  // extern "C"
