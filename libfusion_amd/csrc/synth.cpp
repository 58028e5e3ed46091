// Seeded synthetic RGB-D sequences for the dense-reconstruction path.
//
// Scenes are analytic: an axis-aligned box room seen from inside plus spheres
// and axis-aligned boxes.  Depth is the camera-frame z of the nearest hit with
// Gaussian noise sigma(z) = 0.0012 + 0.0019 (z - 0.4)^2 m and 2% dropouts,
// quantised to uint16 at depth_scale metres per unit.  Colour is a Perlin
// (improved noise, Perlin 2002) albedo per surface under Lambertian shading
// from one point light plus ambient.  All randomness is counter-based
// (SplitMix64 of seed/frame/pixel), so results do not depend on threading.
//
// Camera convention (as libfusion's pinhole model, cuda_imgproc.cu:261):
// x right, y down, z forward; pixel (u, v) sees direction ((u-cx)/fx, (v-cy)/fy, 1).
// Poses are T_world_camera, row-major 4x4 doubles.  World z is up.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

namespace {

inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

inline double u01(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

struct Rng {  // PCG32 (O'Neill 2014), seeded through SplitMix64
    uint64_t state, inc;
    explicit Rng(uint64_t seed, uint64_t stream = 0) {
        inc = (splitmix64(stream ^ 0x5851F42D4C957F2Dull) << 1) | 1u;
        state = 0;
        next();
        state += splitmix64(seed);
        next();
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
        uint32_t rot = (uint32_t)(old >> 59);
        return (xs >> rot) | (xs << ((32 - rot) & 31));
    }
    double uniform() { return ((double)next() + 0.5) / 4294967296.0; }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
};

struct Perlin {  // Ken Perlin's improved noise with a seeded permutation
    int p[512];
    explicit Perlin(uint64_t seed) {
        Rng r(seed, 7);
        int perm[256];
        for (int i = 0; i < 256; ++i) perm[i] = i;
        for (int i = 255; i > 0; --i) {
            int j = (int)(r.next() % (uint32_t)(i + 1));
            std::swap(perm[i], perm[j]);
        }
        for (int i = 0; i < 512; ++i) p[i] = perm[i & 255];
    }
    static double fade(double t) { return t * t * t * (t * (t * 6 - 15) + 10); }
    static double lerp(double t, double a, double b) { return a + t * (b - a); }
    static double grad(int hash, double x, double y, double z) {
        int h = hash & 15;
        double u = h < 8 ? x : y;
        double v = h < 4 ? y : (h == 12 || h == 14 ? x : z);
        return ((h & 1) == 0 ? u : -u) + ((h & 2) == 0 ? v : -v);
    }
    double noise(double x, double y, double z) const {
        double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
        int X = (int)fx & 255, Y = (int)fy & 255, Z = (int)fz & 255;
        x -= fx; y -= fy; z -= fz;
        double u = fade(x), v = fade(y), w = fade(z);
        int A = p[X] + Y, AA = p[A] + Z, AB = p[A + 1] + Z;
        int B = p[X + 1] + Y, BA = p[B] + Z, BB = p[B + 1] + Z;
        return lerp(w, lerp(v, lerp(u, grad(p[AA], x, y, z), grad(p[BA], x - 1, y, z)),
                               lerp(u, grad(p[AB], x, y - 1, z), grad(p[BB], x - 1, y - 1, z))),
                       lerp(v, lerp(u, grad(p[AA + 1], x, y, z - 1), grad(p[BA + 1], x - 1, y, z - 1)),
                               lerp(u, grad(p[AB + 1], x, y - 1, z - 1), grad(p[BB + 1], x - 1, y - 1, z - 1))));
    }
};

struct V3 { double x, y, z; };
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 normalize(V3 a) { double n = std::sqrt(dot(a, a)); return a * (1.0 / n); }

struct Prim {
    int type;       // 0 sphere, 1 box (axis aligned)
    V3 c;           // centre
    V3 h;           // sphere: h.x = radius; box: half extents
    V3 albedo;
    double freq;
};

struct Intr { int w, h; double fx, fy, cx, cy; };

struct Scene {
    V3 room_min, room_max;
    std::vector<V3> room_walls_min, room_walls_max;  // interior partition walls (boxes)
    std::vector<Prim> prims;
    V3 room_albedo[6];
    V3 light;
    Intr K;
    int num_frames;
    int trajectory;   // 0 fixed pose, 1 orbit, 2 random walk
    double orbit_radius, orbit_height, orbit_arc;
    double jitter_t, jitter_r;
    uint64_t seed;
    double depth_scale;
    Perlin perlin;
    std::vector<double> walk;  // random-walk poses, 16 doubles each
    explicit Scene(uint64_t s) : perlin(s) { seed = s; }
};

// Nearest hit of ray o + t d (t > tmin) with the scene.  Returns t, normal, albedo.
bool intersect(const Scene &S, V3 o, V3 d, double &t_hit, V3 &n_hit, V3 &alb, double &freq, int &surf) {
    const double tmin = 1e-4;
    double best = 1e30;
    bool hit = false;
    // Room interior: exit point of the box.
    {
        double tx = d.x > 0 ? (S.room_max.x - o.x) / d.x : (d.x < 0 ? (S.room_min.x - o.x) / d.x : 1e30);
        double ty = d.y > 0 ? (S.room_max.y - o.y) / d.y : (d.y < 0 ? (S.room_min.y - o.y) / d.y : 1e30);
        double tz = d.z > 0 ? (S.room_max.z - o.z) / d.z : (d.z < 0 ? (S.room_min.z - o.z) / d.z : 1e30);
        double t = std::min(tx, std::min(ty, tz));
        if (t > tmin && t < best) {
            best = t; hit = true; freq = 6.0;
            if (t == tx) { n_hit = {d.x > 0 ? -1.0 : 1.0, 0, 0}; surf = d.x > 0 ? 0 : 1; }
            else if (t == ty) { n_hit = {0, d.y > 0 ? -1.0 : 1.0, 0}; surf = d.y > 0 ? 2 : 3; }
            else { n_hit = {0, 0, d.z > 0 ? -1.0 : 1.0}; surf = d.z > 0 ? 4 : 5; }
            alb = S.room_albedo[surf];
        }
    }
    auto box_hit = [&](V3 bmin, V3 bmax, double &t_out, V3 &n_out) -> bool {
        double t0 = -1e30, t1 = 1e30;
        int axis0 = 0; double sgn0 = 1;
        const double oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
        const double lo[3] = {bmin.x, bmin.y, bmin.z}, hi[3] = {bmax.x, bmax.y, bmax.z};
        for (int a = 0; a < 3; ++a) {
            if (std::fabs(dd[a]) < 1e-15) {
                if (oo[a] < lo[a] || oo[a] > hi[a]) return false;
                continue;
            }
            double ta = (lo[a] - oo[a]) / dd[a], tb = (hi[a] - oo[a]) / dd[a];
            double s = -1;
            if (ta > tb) { std::swap(ta, tb); s = 1; }
            if (ta > t0) { t0 = ta; axis0 = a; sgn0 = s; }
            if (tb < t1) t1 = tb;
            if (t0 > t1) return false;
        }
        if (t0 <= tmin) return false;  // camera inside or box behind
        t_out = t0;
        double nn[3] = {0, 0, 0};
        nn[axis0] = sgn0;
        n_out = {nn[0], nn[1], nn[2]};
        return true;
    };
    for (size_t i = 0; i < S.room_walls_min.size(); ++i) {
        double t; V3 n;
        if (box_hit(S.room_walls_min[i], S.room_walls_max[i], t, n) && t < best) {
            best = t; hit = true; n_hit = n; alb = S.room_albedo[(i + 1) % 6]; freq = 6.0; surf = 6;
        }
    }
    for (size_t i = 0; i < S.prims.size(); ++i) {
        const Prim &P = S.prims[i];
        if (P.type == 0) {
            V3 oc = o - P.c;
            double b = dot(oc, d), c = dot(oc, oc) - P.h.x * P.h.x, a = dot(d, d);
            double disc = b * b - a * c;
            if (disc < 0) continue;
            double sq = std::sqrt(disc);
            double t = (-b - sq) / a;
            if (t <= tmin) continue;
            if (t < best) {
                best = t; hit = true;
                n_hit = normalize(o + d * t - P.c);
                alb = P.albedo; freq = P.freq; surf = 7 + (int)i;
            }
        } else {
            double t; V3 n;
            if (box_hit(P.c - P.h, P.c + P.h, t, n) && t < best) {
                best = t; hit = true; n_hit = n; alb = P.albedo; freq = P.freq; surf = 7 + (int)i;
            }
        }
    }
    t_hit = best;
    return hit;
}

double scene_sdf(const Scene &S, V3 p) {
    // Signed distance, positive in free space (inside the room, outside objects).
    auto box_sdf_out = [](V3 q, V3 c, V3 h) {
        double dx = std::fabs(q.x - c.x) - h.x, dy = std::fabs(q.y - c.y) - h.y, dz = std::fabs(q.z - c.z) - h.z;
        double ox = std::max(dx, 0.0), oy = std::max(dy, 0.0), oz = std::max(dz, 0.0);
        return std::sqrt(ox * ox + oy * oy + oz * oz) + std::min(std::max(dx, std::max(dy, dz)), 0.0);
    };
    V3 rc = (S.room_min + S.room_max) * 0.5, rh = (S.room_max - S.room_min) * 0.5;
    double d = -box_sdf_out(p, rc, rh);
    for (size_t i = 0; i < S.room_walls_min.size(); ++i) {
        V3 c = (S.room_walls_min[i] + S.room_walls_max[i]) * 0.5, h = (S.room_walls_max[i] - S.room_walls_min[i]) * 0.5;
        d = std::min(d, box_sdf_out(p, c, h));
    }
    for (const Prim &P : S.prims) {
        double e = P.type == 0 ? std::sqrt(dot(p - P.c, p - P.c)) - P.h.x : box_sdf_out(p, P.c, P.h);
        d = std::min(d, e);
    }
    return d;
}

void look_at(V3 eye, V3 target, double *T) {
    V3 f = normalize(target - eye);
    V3 up = {0, 0, 1};
    V3 r = normalize(cross(f, up));
    V3 dn = cross(f, r);
    double M[16] = {r.x, dn.x, f.x, eye.x, r.y, dn.y, f.y, eye.y, r.z, dn.z, f.z, eye.z, 0, 0, 0, 1};
    std::memcpy(T, M, sizeof(M));
}

void rodrigues(V3 w, double *R) {  // 3x3 row-major
    double th = std::sqrt(dot(w, w));
    double K[9] = {0, -w.z, w.y, w.z, 0, -w.x, -w.y, w.x, 0};
    double A = th < 1e-12 ? 1.0 : std::sin(th) / th;
    double B = th < 1e-12 ? 0.5 : (1 - std::cos(th)) / (th * th);
    double K2[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += K[i * 3 + k] * K[k * 3 + j];
            K2[i * 3 + j] = s;
        }
    for (int i = 0; i < 9; ++i) R[i] = (i % 4 == 0 ? 1.0 : 0.0) + A * K[i] + B * K2[i];
}

void mat_mul4(const double *A, const double *B, double *C) {
    double t[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = 0;
            for (int k = 0; k < 4; ++k) s += A[i * 4 + k] * B[k * 4 + j];
            t[i * 4 + j] = s;
        }
    std::memcpy(C, t, sizeof(t));
}

void jitter(const Scene &S, int frame, double *T) {
    if (S.jitter_t <= 0 && S.jitter_r <= 0) return;
    Rng r(S.seed ^ 0xA5A5A5A5ull, 1000 + (uint64_t)frame);
    // translation uniform in a ball of radius jitter_t, rotation about a uniform axis
    V3 t;
    do { t = {r.uniform(-1, 1), r.uniform(-1, 1), r.uniform(-1, 1)}; } while (dot(t, t) > 1.0);
    t = t * S.jitter_t;
    V3 ax;
    do { ax = {r.uniform(-1, 1), r.uniform(-1, 1), r.uniform(-1, 1)}; } while (dot(ax, ax) > 1.0 || dot(ax, ax) < 1e-6);
    ax = normalize(ax) * (r.uniform() * S.jitter_r);
    double R[9];
    rodrigues(ax, R);
    double J[16] = {R[0], R[1], R[2], t.x, R[3], R[4], R[5], t.y, R[6], R[7], R[8], t.z, 0, 0, 0, 1};
    mat_mul4(T, J, T);
}

void frame_pose(const Scene &S, int frame, double *T) {
    if (S.trajectory == 2) {
        std::memcpy(T, &S.walk[16 * (size_t)frame], 16 * sizeof(double));
        return;
    }
    double th = S.trajectory == 1 ? S.orbit_arc * (double)frame / (double)std::max(1, S.num_frames) : 0.0;
    V3 eye = {S.orbit_radius * std::cos(th), S.orbit_radius * std::sin(th), S.orbit_height};
    V3 ctr = {(S.room_min.x + S.room_max.x) * 0.5, (S.room_min.y + S.room_max.y) * 0.5, S.orbit_height};
    look_at(eye, ctr, T);
    if (S.trajectory == 1) jitter(S, frame, T);
}

V3 random_albedo(Rng &r) { return {r.uniform(0.35, 0.95), r.uniform(0.35, 0.95), r.uniform(0.35, 0.95)}; }

void add_prims(Scene &S, Rng &r, int count, int sphere_only, double r_lo, double r_hi,
               V3 lo, V3 hi, double keep_out_radius) {
    for (int i = 0; i < count; ++i) {
        Prim P;
        P.type = sphere_only ? 0 : (int)(r.next() & 1u);
        for (int tries = 0; tries < 1000; ++tries) {
            double rad = r.uniform(r_lo, r_hi);
            if (P.type == 0) P.h = {rad, rad, rad};
            else P.h = {r.uniform(0.5, 1.0) * rad, r.uniform(0.5, 1.0) * rad, r.uniform(0.5, 1.0) * rad};
            double ext = std::max(P.h.x, std::max(P.h.y, P.h.z)) * (P.type ? 1.7320508 : 1.0);
            P.c = {r.uniform(lo.x, hi.x), r.uniform(lo.y, hi.y), r.uniform(lo.z, hi.z)};
            double rho = std::sqrt(P.c.x * P.c.x + P.c.y * P.c.y);
            if (keep_out_radius <= 0 || rho + ext < keep_out_radius - 0.7) break;
        }
        P.albedo = random_albedo(r);
        P.freq = r.uniform(5.0, 12.0);
        S.prims.push_back(P);
    }
}

Scene *make_scene(int config, uint64_t seed) {
    Scene *S = new Scene(seed);
    Rng r(seed, 1);
    S->light = {0.3, -0.4, 2.6};
    S->depth_scale = 1.0 / 5000.0;
    S->jitter_t = 0; S->jitter_r = 0;
    for (int i = 0; i < 6; ++i) S->room_albedo[i] = random_albedo(r);
    S->orbit_radius = 1.5; S->orbit_height = 1.2; S->orbit_arc = 0;
    if (config == 0) {  // single 320x240 frame, 4x4x3 room + 3 spheres
        S->room_min = {-2, -2, 0}; S->room_max = {2, 2, 3};
        S->K = {320, 240, 262.5, 262.5, 159.5, 119.5};
        S->num_frames = 1; S->trajectory = 0;
        add_prims(*S, r, 3, 1, 0.2, 0.5, {-0.9, -0.9, 0.5}, {0.9, 0.9, 2.2}, 1.5);
    } else if (config == 1 || config == 2) {  // 300-frame orbit at 640x480, 8 primitives
        S->room_min = {-2, -2, 0}; S->room_max = {2, 2, 3};
        S->K = {640, 480, 525.0, 525.0, 319.5, 239.5};
        S->num_frames = 300; S->trajectory = 1; S->orbit_arc = 2.0 * M_PI;
        S->jitter_t = 0.01; S->jitter_r = 0.5 * M_PI / 180.0;
        add_prims(*S, r, 8, 0, 0.2, 0.5, {-0.9, -0.9, 0.3}, {0.9, 0.9, 2.4}, 1.5);
    } else if (config == 3) {  // batched stream: room of 3-6 m sides, size picked by seed
        double sx = r.uniform(3, 6), sy = r.uniform(3, 6), sz = r.uniform(2.5, 3.5);
        S->room_min = {-sx / 2, -sy / 2, 0}; S->room_max = {sx / 2, sy / 2, sz};
        static const int res[4][2] = {{160, 120}, {320, 240}, {640, 480}, {848, 480}};
        int k = (int)(seed % 4);
        double f = 525.0 * res[k][0] / 640.0;
        if (k == 3) f = 525.0 * 480.0 / 480.0;
        S->K = {res[k][0], res[k][1], f, f, res[k][0] / 2.0 - 0.5, res[k][1] / 2.0 - 0.5};
        S->num_frames = 200; S->trajectory = 1; S->orbit_arc = M_PI;
        S->orbit_radius = 0.25 * std::min(sx, sy); S->orbit_height = 1.2;
        S->jitter_t = 0.01; S->jitter_r = 0.5 * M_PI / 180.0;
        double kr = S->orbit_radius;
        add_prims(*S, r, 6, 0, 0.15, 0.4, {-0.6 * kr, -0.6 * kr, 0.3}, {0.6 * kr, 0.6 * kr, 2.0}, kr);
    } else {  // config 4: 8x8x3 multi-room, 40 primitives, random walk
        S->room_min = {-4, -4, 0}; S->room_max = {4, 4, 3};
        S->room_walls_min.push_back({-0.05, -4, 0}); S->room_walls_max.push_back({0.05, -1.0, 3});
        S->room_walls_min.push_back({-0.05, 1.0, 0}); S->room_walls_max.push_back({0.05, 4, 3});
        S->room_walls_min.push_back({-4, -0.05, 0}); S->room_walls_max.push_back({-1.0, 0.05, 3});
        S->room_walls_min.push_back({1.0, -0.05, 0}); S->room_walls_max.push_back({4, 0.05, 3});
        S->K = {640, 480, 525.0, 525.0, 319.5, 239.5};
        S->num_frames = 1000; S->trajectory = 2;
        add_prims(*S, r, 40, 0, 0.15, 0.4, {-3.6, -3.6, 0.2}, {3.6, 3.6, 2.5}, 0);
        // random walk: 1.5 cm / 1 degree steps, kept inside the room
        S->walk.resize(16 * (size_t)S->num_frames);
        double T[16];
        look_at({-2.0, -2.0, 1.2}, {-1.0, -2.0, 1.2}, T);
        Rng w(seed, 99);
        for (int f = 0; f < S->num_frames; ++f) {
            std::memcpy(&S->walk[16 * (size_t)f], T, sizeof(T));
            V3 ax;
            do { ax = {w.uniform(-1, 1), w.uniform(-1, 1), w.uniform(-1, 1)}; } while (dot(ax, ax) > 1 || dot(ax, ax) < 1e-6);
            ax = normalize(ax) * (M_PI / 180.0);
            ax.x *= 0.3; ax.z *= 0.3;  // mostly yaw (camera y axis is down)
            double R[9];
            rodrigues(ax, R);
            double step[16] = {R[0], R[1], R[2], 0, R[3], R[4], R[5], 0, R[6], R[7], R[8], 0.015, 0, 0, 0, 1};
            double N[16];
            mat_mul4(T, step, N);
            // stay inside [-3.5, 3.5]^2 and away from partition walls: reflect heading
            if (std::fabs(N[3]) > 3.5 || std::fabs(N[7]) > 3.5 || std::fabs(N[3]) < 0.35 || std::fabs(N[7]) < 0.35) {
                double turn[16];
                double Ry[9];
                rodrigues({0, M_PI * 0.75, 0}, Ry);
                double tt[16] = {Ry[0], Ry[1], Ry[2], 0, Ry[3], Ry[4], Ry[5], 0, Ry[6], Ry[7], Ry[8], 0, 0, 0, 0, 1};
                std::memcpy(turn, tt, sizeof(tt));
                mat_mul4(T, turn, N);
            }
            N[11] = 1.2;  // keep height
            std::memcpy(T, N, sizeof(T));
            // re-orthonormalise rotation (keep z-up roll-free look)
            V3 f3 = normalize({T[2], T[6], 0.0});
            look_at({T[3], T[7], T[11]}, {T[3] + f3.x, T[7] + f3.y, T[11]}, T);
        }
    }
    return S;
}

}  // namespace

extern "C" {

void *synth_scene_create(int config, unsigned long long seed) { return make_scene(config, seed); }
void synth_scene_free(void *h) { delete static_cast<Scene *>(h); }
int synth_num_frames(void *h) { return static_cast<Scene *>(h)->num_frames; }
void synth_intrinsics(void *h, int *wh, double *f4) {
    Scene *S = static_cast<Scene *>(h);
    wh[0] = S->K.w; wh[1] = S->K.h;
    f4[0] = S->K.fx; f4[1] = S->K.fy; f4[2] = S->K.cx; f4[3] = S->K.cy;
}
double synth_depth_scale(void *h) { return static_cast<Scene *>(h)->depth_scale; }
void synth_pose(void *h, int frame, double *T16) { frame_pose(*static_cast<Scene *>(h), frame, T16); }

// Render one view.  T16: T_world_camera (row-major 4x4); intr: w, h as ints in
// wh and fx, fy, cx, cy in f4.  noise != 0 adds depth noise and dropouts using
// frame_id as the counter.  depth: uint16 [h*w]; rgb: uint8 [h*w*3] (may be null);
// depth_true: float metres without noise (may be null).
void synth_render(void *h, const double *T16, const int *wh, const double *f4, int noise,
                  int frame_id, uint16_t *depth, uint8_t *rgb, float *depth_true) {
    const Scene &S = *static_cast<Scene *>(h);
    const int W = wh[0], H = wh[1];
    const double fx = f4[0], fy = f4[1], cx = f4[2], cy = f4[3];
    const V3 o = {T16[3], T16[7], T16[11]};
#pragma omp parallel for schedule(dynamic, 4)
    for (int v = 0; v < H; ++v) {
        for (int u = 0; u < W; ++u) {
            const size_t idx = (size_t)v * W + u;
            V3 dc = {(u - cx) / fx, (v - cy) / fy, 1.0};
            V3 dw = {T16[0] * dc.x + T16[1] * dc.y + T16[2] * dc.z,
                     T16[4] * dc.x + T16[5] * dc.y + T16[6] * dc.z,
                     T16[8] * dc.x + T16[9] * dc.y + T16[10] * dc.z};
            double t; V3 n, alb; double freq; int surf = 0;
            bool hit = intersect(S, o, dw, t, n, alb, freq, surf);
            double z = hit ? t : 0.0;  // dc.z == 1 so t is the camera-frame depth
            if (depth_true) depth_true[idx] = (float)z;
            double zn = z;
            if (noise && hit) {
                uint64_t key = splitmix64(S.seed ^ 0xD1B54A32D192ED03ull) ^ ((uint64_t)frame_id << 32) ^ (uint64_t)idx;
                uint64_t h1 = splitmix64(key), h2 = splitmix64(h1), h3 = splitmix64(h2);
                double sigma = 0.0012 + 0.0019 * (z - 0.4) * (z - 0.4);
                double g = std::sqrt(-2.0 * std::log(std::max(u01(h1), 1e-300))) * std::cos(2.0 * M_PI * u01(h2));
                zn = z + sigma * g;
                if (u01(h3) < 0.02) zn = 0.0;
            }
            double q = std::floor(zn / S.depth_scale + 0.5);
            depth[idx] = (uint16_t)std::min(65535.0, std::max(0.0, q));
            if (rgb) {
                double c[3] = {0, 0, 0};
                if (hit) {
                    V3 p = o + dw * t;
                    double tex = 0.0, amp = 1.0, fr = freq;
                    for (int oc = 0; oc < 3; ++oc) {
                        tex += amp * S.perlin.noise(p.x * fr + 17.0 * surf, p.y * fr + 3.0 * surf, p.z * fr);
                        amp *= 0.5; fr *= 2.0;
                    }
                    double m = std::min(1.0, std::max(0.0, 0.55 + 0.6 * tex));
                    V3 l = normalize(S.light - p);
                    if (dot(n, dw) > 0) n = n * -1.0;
                    double shade = 0.3 + 0.7 * std::max(0.0, dot(n, l));
                    c[0] = alb.x * m * shade; c[1] = alb.y * m * shade; c[2] = alb.z * m * shade;
                }
                for (int k = 0; k < 3; ++k)
                    rgb[3 * idx + k] = (uint8_t)std::min(255.0, std::max(0.0, std::floor(255.0 * c[k] + 0.5)));
            }
        }
    }
}

// Render frame `frame` of the scene's own trajectory with noise.
void synth_render_frame(void *h, int frame, uint16_t *depth, uint8_t *rgb) {
    Scene &S = *static_cast<Scene *>(h);
    double T[16];
    frame_pose(S, frame, T);
    int wh[2] = {S.K.w, S.K.h};
    double f4[4] = {S.K.fx, S.K.fy, S.K.cx, S.K.cy};
    synth_render(h, T, wh, f4, 1, frame, depth, rgb, nullptr);
}

// Signed distance to the analytic surface at n points (xyz triples).
void synth_sdf(void *h, const double *pts, int n, double *out) {
    const Scene &S = *static_cast<Scene *>(h);
    for (int i = 0; i < n; ++i) out[i] = scene_sdf(S, {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]});
}

}  // extern "C"
