// mesh.cpp -- procedural workloads: displaced icospheres / geodesic spheres and the pinhole camera.
//
// The reference renders glTF scenes (src/base/scene.cpp); the benchmark workloads here are
// procedural instead: an icosphere subdivided s times (20*4^s triangles) or a geodesic sphere of
// frequency k (20*k^2 triangles), whose vertices are displaced radially by seeded 3-octave value
// noise (lacunarity 2, gain 0.5, values in [-1, 1]).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "host_common.h"

namespace nbvh {
namespace {

inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Lattice value in [-1, 1] at integer point (i, j, k).
inline double lattice(int64_t i, int64_t j, int64_t k, uint64_t seed)
{
    uint64_t h = mix64(seed ^ mix64((uint64_t)i ^ mix64((uint64_t)j ^ mix64((uint64_t)k))));
    return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

inline double fade(double t) { return t * t * (3.0 - 2.0 * t); }

double value_noise(double x, double y, double z, uint64_t seed)
{
    const double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    const int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
    const double tx = fade(x - fx), ty = fade(y - fy), tz = fade(z - fz);
    double c[2][2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int d = 0; d < 2; d++) c[a][b][d] = lattice(ix + a, iy + b, iz + d, seed);
    double r = 0.0;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int d = 0; d < 2; d++)
                r += c[a][b][d] * (a ? tx : 1.0 - tx) * (b ? ty : 1.0 - ty) * (d ? tz : 1.0 - tz);
    return r;
}

// 3 octaves, base frequency 2, lacunarity 2, gain 0.5, normalised to [-1, 1].
double fbm3(double x, double y, double z, uint64_t seed)
{
    double sum = 0.0, amp = 1.0, freq = 2.0, norm = 0.0;
    for (int o = 0; o < 3; o++) {
        sum += amp * value_noise(x * freq, y * freq, z * freq, seed + (uint64_t)o * 0x632BE59BD9B4E019ull);
        norm += amp;
        amp *= 0.5;
        freq *= 2.0;
    }
    return sum / norm;
}

// Unit icosahedron: 12 vertices, 20 counter-clockwise faces.
void icosahedron(std::vector<double>& v, std::vector<uint32_t>& f)
{
    const double phi = (1.0 + std::sqrt(5.0)) * 0.5;
    v = {-1, phi, 0, 1, phi, 0, -1, -phi, 0, 1, -phi, 0, 0, -1, phi, 0, 1, phi,
         0, -1, -phi, 0, 1, -phi, phi, 0, -1, phi, 0, 1, -phi, 0, -1, -phi, 0, 1};
    for (size_t i = 0; i < v.size(); i += 3) {
        const double n = std::sqrt(v[i] * v[i] + v[i + 1] * v[i + 1] + v[i + 2] * v[i + 2]);
        v[i] /= n;
        v[i + 1] /= n;
        v[i + 2] /= n;
    }
    f = {0, 11, 5, 0, 5,  1, 0, 1, 7, 0, 7,  10, 0, 10, 11, 1, 5, 9, 5, 11,
         4, 11, 10, 2, 10, 7, 6, 7, 1, 8, 3, 9,  4,  3, 4,  2, 3, 2, 6, 3,
         6, 8,  3,  8, 9,  4, 9, 5, 2, 4, 11, 6, 2,  10, 8, 6, 7, 9, 8, 1};
}

// Unit-sphere directions v (xyz triples) -> displaced, scaled, translated float vertices.
void displace(const std::vector<double>& v, uint64_t seed, float amplitude, const float center[3], float radius,
              std::vector<float>& verts)
{
    verts.resize(v.size());
    for (size_t i = 0; i < v.size(); i += 3) {
        const double n = fbm3(v[i], v[i + 1], v[i + 2], seed);
        const double r = (double)radius * (1.0 + (double)amplitude * n);
        for (int c = 0; c < 3; c++) verts[i + c] = (float)(v[i + c] * r + (double)center[c]);
    }
}

}  // namespace

// Geodesic sphere of frequency k: every icosahedron face split into a k x k triangular grid
// (20 k^2 triangles, 10 k^2 + 2 vertices), grid points projected to the unit sphere, then
// displaced like the icosphere. Vertices on shared edges are shared.
void geosphere(uint32_t k, uint64_t seed, float amplitude, const float center[3], float radius,
               std::vector<float>& verts, std::vector<uint32_t>& tris)
{
    std::vector<double> base;
    std::vector<uint32_t> f;
    icosahedron(base, f);
    // Edge ids (unordered corner pairs), k - 1 interior points each, stored from the lower
    // corner index to the higher one.
    std::unordered_map<uint64_t, uint32_t> edge_id;
    for (size_t t = 0; t < f.size(); t += 3)
        for (int e = 0; e < 3; e++) {
            const uint32_t a = f[t + e], b = f[t + (e + 1) % 3];
            const uint64_t key = a < b ? ((uint64_t)a << 32) | b : ((uint64_t)b << 32) | a;
            edge_id.emplace(key, (uint32_t)edge_id.size());
        }
    const uint64_t n_edge_pts = 30ull * (k - 1), n_face_pts = k >= 2 ? (uint64_t)(k - 1) * (k - 2) / 2 : 0;
    const uint64_t n_verts = 12 + n_edge_pts + 20 * n_face_pts;
    std::vector<double> v(3 * n_verts);
    auto lerp_dir = [&](uint64_t out, const double* a, const double* b, const double* c, double u, double w) {
        double p[3];
        for (int d = 0; d < 3; d++) p[d] = a[d] + (b[d] - a[d]) * u + (c[d] - a[d]) * w;
        const double n = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        for (int d = 0; d < 3; d++) v[3 * out + d] = p[d] / n;
    };
    for (int i = 0; i < 36; i++) v[i] = base[i];
    for (const auto& e : edge_id) {
        const uint32_t a = (uint32_t)(e.first >> 32), b = (uint32_t)e.first;
        for (uint32_t t = 1; t < k; t++)
            lerp_dir(12 + (uint64_t)e.second * (k - 1) + (t - 1), &base[3 * a], &base[3 * b], &base[3 * a],
                     (double)t / k, 0.0);
    }
    // Global index of grid point (i, j) of face fi: P = A + (B - A) i/k + (C - A) j/k.
    auto edge_point = [&](uint32_t a, uint32_t b, uint32_t t) -> uint32_t {
        const uint64_t key = a < b ? ((uint64_t)a << 32) | b : ((uint64_t)b << 32) | a;
        const uint32_t id = edge_id.at(key);
        return (uint32_t)(12 + (uint64_t)id * (k - 1) + (a < b ? t - 1 : k - t - 1));
    };
    auto vid = [&](size_t fi, uint32_t i, uint32_t j) -> uint32_t {
        const uint32_t A = f[3 * fi], B = f[3 * fi + 1], C = f[3 * fi + 2];
        if (j == 0) return i == 0 ? A : (i == k ? B : edge_point(A, B, i));
        if (i == 0) return j == k ? C : edge_point(A, C, j);
        if (i + j == k) return edge_point(B, C, j);
        const uint64_t row_start = (uint64_t)(j - 1) * (2 * k - j - 2) / 2;  // interior points of rows 1..j-1
        return (uint32_t)(12 + n_edge_pts + fi * n_face_pts + row_start + (i - 1));
    };
    for (size_t fi = 0; fi < 20; fi++) {
        const double *A = &base[3 * f[3 * fi]], *B = &base[3 * f[3 * fi + 1]], *C = &base[3 * f[3 * fi + 2]];
        for (uint32_t j = 1; j + 1 < k; j++)
            for (uint32_t i = 1; i + j < k; i++) lerp_dir(vid(fi, i, j), A, B, C, (double)i / k, (double)j / k);
    }
    tris.clear();
    tris.reserve(60ull * k * k);
    for (size_t fi = 0; fi < 20; fi++)
        for (uint32_t j = 0; j < k; j++)
            for (uint32_t i = 0; i + j < k; i++) {
                const uint32_t p00 = vid(fi, i, j), p10 = vid(fi, i + 1, j), p01 = vid(fi, i, j + 1);
                tris.insert(tris.end(), {p00, p10, p01});
                if (i + j + 1 < k) {
                    const uint32_t p11 = vid(fi, i + 1, j + 1);
                    tris.insert(tris.end(), {p10, p11, p01});
                }
            }
    displace(v, seed, amplitude, center, radius, verts);
}

void icosphere(uint32_t subdivisions, uint64_t seed, float amplitude, const float center[3], float radius,
               std::vector<float>& verts, std::vector<uint32_t>& tris)
{
    std::vector<double> v;
    std::vector<uint32_t> f;
    icosahedron(v, f);
    for (uint32_t s = 0; s < subdivisions; s++) {
        std::unordered_map<uint64_t, uint32_t> mid;
        mid.reserve(f.size());
        auto midpoint = [&](uint32_t a, uint32_t b) -> uint32_t {
            const uint64_t key = a < b ? ((uint64_t)a << 32) | b : ((uint64_t)b << 32) | a;
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            double m[3];
            for (int c = 0; c < 3; c++) m[c] = 0.5 * (v[3 * a + c] + v[3 * b + c]);
            const double n = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
            const uint32_t idx = (uint32_t)(v.size() / 3);
            for (int c = 0; c < 3; c++) v.push_back(m[c] / n);
            mid.emplace(key, idx);
            return idx;
        };
        std::vector<uint32_t> g;
        g.reserve(f.size() * 4);
        for (size_t t = 0; t < f.size(); t += 3) {
            const uint32_t a = f[t], b = f[t + 1], c = f[t + 2];
            const uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            const uint32_t nt[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            g.insert(g.end(), nt, nt + 12);
        }
        f.swap(g);
    }
    displace(v, seed, amplitude, center, radius, verts);
    tris = f;
}

// CudaBackend::update_camera (src/cuda/cuda_backend.cpp:595-611), float arithmetic.
void camera_look_at(const float eye[3], const float target[3], const float up[3], float fovy_deg, uint32_t w,
                    uint32_t h, nbvh_camera& cam)
{
    auto normalize = [](float v[3]) {
        const float inv = 1.f / std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
        for (int c = 0; c < 3; c++) v[c] *= inv;
    };
    auto cross = [](const float a[3], const float b[3], float o[3]) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    const float plane_y = 2.f * std::tan(0.5f * fovy_deg * 0.01745329251994329577f);
    const float plane_x = plane_y * ((float)w / (float)h);
    float dir[3] = {target[0] - eye[0], target[1] - eye[1], target[2] - eye[2]};
    normalize(dir);
    float du[3], dv[3], tmp[3];
    cross(dir, up, du);
    normalize(du);
    for (int c = 0; c < 3; c++) du[c] *= plane_x;
    cross(du, dir, tmp);
    normalize(tmp);
    for (int c = 0; c < 3; c++) dv[c] = -tmp[c] * plane_y;
    for (int c = 0; c < 3; c++) {
        cam.origin[c]       = eye[c];
        cam.du[c]           = du[c];
        cam.dv[c]           = dv[c];
        cam.dir_top_left[c] = dir[c] - 0.5f * du[c] - 0.5f * dv[c];
    }
}

}  // namespace nbvh
