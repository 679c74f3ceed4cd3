// Deterministic procedural scene generators for the benchmark workloads:
// displaced icospheres, random boxes and vertex normals.
#include <math.h>
#include <stdint.h>

#include <unordered_map>
#include <vector>

#include "../../include/ilm_gpu.h"
#include "internal.hpp"

namespace {

// splitmix64 stream (Steele, Lea, Flood 2014)
struct SplitMix {
    uint64_t s;
    explicit SplitMix(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

uint32_t hash3(int32_t x, int32_t y, int32_t z, uint32_t seed) {
    uint64_t h = (uint64_t)(uint32_t)x * 0x8DA6B343ull ^ (uint64_t)(uint32_t)y * 0xD8163841ull ^
                 (uint64_t)(uint32_t)z * 0xCB1AB31Full ^ ((uint64_t)seed << 32);
    h ^= h >> 33;
    h *= 0xFF51AFD7ED558CCDull;
    h ^= h >> 33;
    h *= 0xC4CEB9FE1A85EC53ull;
    h ^= h >> 33;
    return (uint32_t)h;
}

double lattice(int32_t x, int32_t y, int32_t z, uint32_t seed) {
    return (double)hash3(x, y, z, seed) * (2.0 / 4294967295.0) - 1.0;  // [-1, 1]
}

double fade(double t) { return t * t * t * (t * (t * 6 - 15) + 10); }

// Trilinear value noise with quintic fade, in [-1, 1].
double value_noise(double x, double y, double z, uint32_t seed) {
    double fx = floor(x), fy = floor(y), fz = floor(z);
    int32_t ix = (int32_t)fx, iy = (int32_t)fy, iz = (int32_t)fz;
    double tx = fade(x - fx), ty = fade(y - fy), tz = fade(z - fz);
    double c[2][2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int d = 0; d < 2; ++d) c[a][b][d] = lattice(ix + a, iy + b, iz + d, seed);
    double x00 = c[0][0][0] + (c[1][0][0] - c[0][0][0]) * tx, x10 = c[0][1][0] + (c[1][1][0] - c[0][1][0]) * tx;
    double x01 = c[0][0][1] + (c[1][0][1] - c[0][0][1]) * tx, x11 = c[0][1][1] + (c[1][1][1] - c[0][1][1]) * tx;
    double y0 = x00 + (x10 - x00) * ty, y1 = x01 + (x11 - x01) * ty;
    return y0 + (y1 - y0) * tz;
}

// Fractal sum: base frequency 4, lacunarity 2, gain 0.5, normalised to [-1, 1].
double fbm(double x, double y, double z, int octaves, uint32_t seed) {
    double sum = 0, amp = 1, norm = 0, f = 4.0;
    for (int o = 0; o < octaves; ++o) {
        sum += amp * value_noise(x * f, y * f, z * f, seed + 1013u * (uint32_t)o);
        norm += amp;
        amp *= 0.5;
        f *= 2.0;
    }
    return norm > 0 ? sum / norm : 0.0;
}

}  // namespace

extern "C" {

int ilm_scenegen_icosphere_size(int level, int64_t* nv, int64_t* nt) {
    if (level < 0 || level > 11 || !nv || !nt) return ilm::fail(ILM_ERR_INVALID, "icosphere level must be 0..11");
    int64_t f = 20;
    for (int i = 0; i < level; ++i) f *= 4;
    *nt = f;
    *nv = f / 2 + 2;
    return ILM_OK;
}

int ilm_scenegen_icosphere(int level, float amplitude, int octaves, uint32_t seed, float* pos, int32_t* idx) {
    int64_t nv, nt;
    int rc = ilm_scenegen_icosphere_size(level, &nv, &nt);
    if (rc) return rc;
    if (!pos || !idx) return ilm::fail(ILM_ERR_INVALID, "null output");
    const double t = (1.0 + sqrt(5.0)) / 2.0;
    std::vector<double> v = {-1, t, 0, 1, t, 0, -1, -t, 0, 1, -t, 0, 0, -1, t, 0, 1, t,
                             0, -1, -t, 0, 1, -t, t, 0, -1, t, 0, 1, -t, 0, -1, -t, 0, 1};
    for (int i = 0; i < 12; ++i) {
        double l = sqrt(v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
        for (int k = 0; k < 3; ++k) v[3 * i + k] /= l;
    }
    // faces wound counter-clockwise seen from outside
    std::vector<int32_t> f = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4,
                              11, 10, 2, 10, 7, 6, 7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8,
                              3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    v.reserve(3 * nv);
    for (int lv = 0; lv < level; ++lv) {
        std::unordered_map<uint64_t, int32_t> mid;
        mid.reserve(f.size());
        std::vector<int32_t> nf;
        nf.reserve(f.size() * 4);
        auto midpoint = [&](int32_t a, int32_t b) -> int32_t {
            uint64_t key = ((uint64_t)(uint32_t)std::min(a, b) << 32) | (uint32_t)std::max(a, b);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            double m[3];
            for (int k = 0; k < 3; ++k) m[k] = v[3 * a + k] + v[3 * b + k];
            double l = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
            int32_t id = (int32_t)(v.size() / 3);
            for (int k = 0; k < 3; ++k) v.push_back(m[k] / l);
            mid.emplace(key, id);
            return id;
        };
        for (size_t i = 0; i < f.size(); i += 3) {
            int32_t a = f[i], b = f[i + 1], c = f[i + 2];
            int32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            int32_t q[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            nf.insert(nf.end(), q, q + 12);
        }
        f.swap(nf);
    }
    if ((int64_t)v.size() != 3 * nv || (int64_t)f.size() != 3 * nt)
        return ilm::fail(ILM_ERR_STATE, "icosphere size mismatch");
    for (int64_t i = 0; i < nv; ++i) {
        double x = v[3 * i], y = v[3 * i + 1], z = v[3 * i + 2];
        double r = 1.0 + amplitude * (octaves > 0 ? fbm(x, y, z, octaves, seed) : 0.0);
        pos[3 * i] = (float)(x * r);
        pos[3 * i + 1] = (float)(y * r);
        pos[3 * i + 2] = (float)(z * r);
    }
    for (int64_t i = 0; i < 3 * nt; ++i) idx[i] = f[i];
    return ILM_OK;
}

int ilm_scenegen_boxes(int n, float emin, float emax, float radius, uint32_t seed, float* pos, int32_t* idx,
                       int64_t vertex_offset) {
    if (n < 0 || !pos || !idx) return ilm::fail(ILM_ERR_INVALID, "bad box arguments");
    SplitMix rng(0xB0C5ull ^ ((uint64_t)seed << 16));
    // unit cube corners, faces as (corner quads) wound counter-clockwise from outside
    static const int faces[6][4] = {{0, 2, 3, 1}, {4, 5, 7, 6}, {0, 1, 5, 4}, {2, 6, 7, 3}, {0, 4, 6, 2}, {1, 3, 7, 5}};
    for (int b = 0; b < n; ++b) {
        double e[3];
        for (int k = 0; k < 3; ++k) e[k] = emin + (emax - emin) * rng.uniform();
        // uniform rotation (Shoemake's quaternion)
        double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
        double qx = sqrt(1 - u1) * sin(2 * M_PI * u2), qy = sqrt(1 - u1) * cos(2 * M_PI * u2);
        double qz = sqrt(u1) * sin(2 * M_PI * u3), qw = sqrt(u1) * cos(2 * M_PI * u3);
        double R[3][3] = {{1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)},
                          {2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)},
                          {2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)}};
        // centre uniform in the ball
        double c[3];
        for (;;) {
            for (int k = 0; k < 3; ++k) c[k] = (2 * rng.uniform() - 1) * radius;
            if (c[0] * c[0] + c[1] * c[1] + c[2] * c[2] <= (double)radius * radius) break;
        }
        double corner[8][3];
        for (int i = 0; i < 8; ++i) {
            double l[3] = {((i & 1) ? 0.5 : -0.5) * e[0], ((i & 2) ? 0.5 : -0.5) * e[1], ((i & 4) ? 0.5 : -0.5) * e[2]};
            for (int r = 0; r < 3; ++r) corner[i][r] = c[r] + R[r][0] * l[0] + R[r][1] * l[1] + R[r][2] * l[2];
        }
        float* p = pos + (int64_t)b * 24 * 3;
        int32_t* t = idx + (int64_t)b * 12 * 3;
        for (int fc = 0; fc < 6; ++fc) {
            for (int k = 0; k < 4; ++k)
                for (int r = 0; r < 3; ++r) p[(fc * 4 + k) * 3 + r] = (float)corner[faces[fc][k]][r];
            int32_t base = (int32_t)(vertex_offset + (int64_t)b * 24 + fc * 4);
            int32_t q[6] = {base, base + 1, base + 2, base, base + 2, base + 3};
            for (int k = 0; k < 6; ++k) t[fc * 6 + k] = q[k];
        }
    }
    return ILM_OK;
}

int ilm_scenegen_vertex_normals(const float* pos, int64_t nv, const int32_t* idx, int64_t nt, float* nrm) {
    if (!pos || !idx || !nrm) return ilm::fail(ILM_ERR_INVALID, "null argument");
    std::vector<double> acc(3 * nv, 0.0);
    for (int64_t i = 0; i < nt; ++i) {
        const float* a = pos + 3 * (int64_t)idx[3 * i];
        const float* b = pos + 3 * (int64_t)idx[3 * i + 1];
        const float* c = pos + 3 * (int64_t)idx[3 * i + 2];
        double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        for (int v = 0; v < 3; ++v)
            for (int k = 0; k < 3; ++k) acc[3 * (int64_t)idx[3 * i + v] + k] += n[k];
    }
    for (int64_t i = 0; i < nv; ++i) {
        double l = sqrt(acc[3 * i] * acc[3 * i] + acc[3 * i + 1] * acc[3 * i + 1] + acc[3 * i + 2] * acc[3 * i + 2]);
        for (int k = 0; k < 3; ++k) nrm[3 * i + k] = l > 0 ? (float)(acc[3 * i + k] / l) : 0.0f;
    }
    return ILM_OK;
}

}  // extern "C"
