// cgs_o1.cpp — "O1": C++ restatement of the reference's Rayon CPU render
// path, used ONLY by tests/ (pinned by the reference's golden image hash) and
// by the cpu_baseline leg of bench.py, where it is the CPU path timed beside
// the GPU.  Nothing in the product links or calls it.
//
// What it restates (same floating-point operation order as the Rust source,
// compiled with -ffp-contract=off; the elementary functions are the C
// library's, as Rust's f32::{exp, sin, cos, acos, cbrt} call them):
//   render_scene_cpu                 src/render/mod.rs:288-305
//   project_scene_cpu (chunks 2048)  src/render/mod.rs:140-212
//   project_ewa_splat + helpers      src/math/projection.rs:59-172, 243-360
//     (glam's Mat4::transform_point3 and Mat3 * Mat3 * Mat3^T as glam 0.28
//      evaluates them: per column, x*a + y*b then + z*c)
//   bin_and_sort_gaussians_cpu       src/render/tiles/cpu_reference.rs:66-302
//     3-sigma square bounds          src/render/tiles/mod.rs:97-236
//     SortEntry order (tile, depth bits, gaussian id)
//                                    src/render/tiles/keys.rs:8-95
//   rasterize_scene_cpu              src/render/rasterizer/cpu_reference.rs:120-240
//   evaluate_sh                      src/render/rasterizer/sh.rs:5-137
//   RasterizedImage::from_pixels / pixel_hash (FNV-1a)
//                                    src/render/rasterizer/cpu_reference.rs:53-112
//   fibonacci_cluster                src/scene/synthetic.rs:6-42
//
// The parallel structure follows src/parallel.rs: projection in parallel
// 2048-splat chunks, parallel tile counts, a serial prefix sum, a parallel
// scatter in which every entry finds its splat by binary search over the
// offsets (cpu_reference.rs:168-205), a parallel sort (par_sort_unstable ->
// libstdc++'s OpenMP multiway merge sort; the order is total, so the result
// does not depend on the algorithm), parallel per-tile range searches and a
// parallel per-pixel compositor.  OpenMP threads stand in for Rayon's pool.
//
// Pinned by: the golden image hash 0xbdd6_3815_3115_0f06 of
// tests/gpu_parity_tests.rs:293-312 (fibonacci_cluster(256, 0.5, [0,0,2]),
// identity camera f=180, 160x120, SH degree 2), the known answers of
// projection.rs:556-740 and cpu_reference.rs:247-358, and thread-count
// independence of the image.
#include <stdint.h>
#include <string.h>
#include <math.h>
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <parallel/algorithm>
#include <vector>

#define O1_API extern "C" __attribute__((visibility("default")))

namespace {

constexpr float LOW_PASS = 0.3f;          // projection.rs:4
constexpr float MIN_DET = 1e-6f;          // projection.rs:7
constexpr float MIN_RADIUS = 1.0f;        // projection.rs:13
constexpr float MAX_RADIUS = 1024.0f;     // projection.rs:16
constexpr uint32_t TILE = 16;             // tiles/mod.rs:10
constexpr size_t PROJECT_CHUNK = 2048;    // render/mod.rs:140

// Rust f32::clamp (NaN passes through), f32::max / f32::min (NaN-ignoring).
inline float rclamp(float v, float lo, float hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}
inline float rmax(float a, float b) { return fmaxf(a, b); }
inline float rmin(float a, float b) { return fminf(a, b); }

// Rust `f as i32`: saturating, NaN -> 0.
inline int32_t sat_i32(float v) {
    if (v != v) return 0;
    if (v >= 2147483648.0f) return INT32_MAX;
    if (v <= -2147483648.0f) return INT32_MIN;
    return (int32_t)v;
}

struct V3 {
    float x, y, z;
};

// Camera block (CameraUniforms, camera.rs:18-38): view[16] column-major,
// fx, fy, cx, cy, viewport w, h, near, pad.
struct Cam {
    float m[16];
    float fx, fy, cx, cy, vw, vh, near_plane;
};

inline Cam load_cam(const float* c) {
    Cam k;
    memcpy(k.m, c, sizeof(k.m));
    k.fx = c[16];
    k.fy = c[17];
    k.cx = c[18];
    k.cy = c[19];
    k.vw = c[20];
    k.vh = c[21];
    k.near_plane = c[22];
    return k;
}

struct Splat {
    float px, py, depth;
    int32_t radius;
    float conic[3];
};

// projection.rs:278-360 (project_ewa_splat)
Splat project_ewa_splat(V3 p, const float* cov, const Cam& k) {
    const float* m = k.m;
    // Mat4::transform_point3: x_axis*p.x, + y_axis*p.y, + z_axis*p.z, + w_axis
    const float tx = ((m[0] * p.x + m[4] * p.y) + m[8] * p.z) + m[12];
    const float ty = ((m[1] * p.x + m[5] * p.y) + m[9] * p.z) + m[13];
    const float tz = ((m[2] * p.x + m[6] * p.y) + m[10] * p.z) + m[14];
    Splat s{0.0f, 0.0f, tz, 0, {0.0f, 0.0f, 0.0f}};
    if (tz <= k.near_plane) return s;

    // compute_jacobian (59-68)
    const float inv_tz = 1.0f / tz;
    const float inv_tz2 = inv_tz * inv_tz;
    const float j00 = k.fx * inv_tz;
    const float j02 = -k.fx * tx * inv_tz2;
    const float j11 = k.fy * inv_tz;
    const float j12 = -k.fy * ty * inv_tz2;

    // transform_covariance_to_cam (74-92): W * Sigma * W^T, W = upper-left 3x3
    // columns of the view matrix, each Mat3 product evaluated per column.
    const float W[3][3] = {{m[0], m[1], m[2]}, {m[4], m[5], m[6]}, {m[8], m[9], m[10]}};
    const float S[3][3] = {{cov[0], cov[1], cov[2]}, {cov[1], cov[3], cov[4]}, {cov[2], cov[4], cov[5]}};
    float M[3][3];  // M column j = W * S column j
    for (int j = 0; j < 3; ++j)
        for (int r = 0; r < 3; ++r)
            M[j][r] = (W[0][r] * S[j][0] + W[1][r] * S[j][1]) + W[2][r] * S[j][2];
    float R[3][3];  // R column j = M * (W^T column j) = M * (row j of W)
    for (int j = 0; j < 3; ++j)
        for (int r = 0; r < 3; ++r)
            R[j][r] = (M[0][r] * W[0][j] + M[1][r] * W[1][j]) + M[2][r] * W[2][j];
    const float c_xx = R[0][0], c_xy = R[0][1], c_xz = R[0][2];
    const float c_yy = R[1][1], c_yz = R[1][2], c_zz = R[2][2];

    // project_covariance_2d (99-130)
    const float m00 = j00 * c_xx + j02 * c_xz;
    const float m01 = j00 * c_xy + j02 * c_yz;
    const float m02 = j00 * c_xz + j02 * c_zz;
    const float m11 = j11 * c_yy + j12 * c_yz;
    const float m12 = j11 * c_yz + j12 * c_zz;
    const float sxx = m00 * j00 + m02 * j02;
    const float sxy = m01 * j11 + m02 * j12;
    const float syy = m11 * j11 + m12 * j12;

    // compute_conic_and_radius (138-172)
    const float a = sxx + LOW_PASS;
    const float b = sxy;
    const float c = syy + LOW_PASS;
    const float det = a * c - b * b;
    if (det <= MIN_DET) return s;
    const float inv_det = 1.0f / det;
    const float conic0 = c * inv_det, conic1 = -b * inv_det, conic2 = a * inv_det;
    const float mid = 0.5f * (a + c);
    const float discr = sqrtf(rmax(mid * mid - det, 0.1f));
    const float l1 = mid + discr;
    const float l2 = rmax(mid - discr, 0.1f);
    const float lmax = rmax(l1, l2);
    const float radius = rclamp(ceilf(3.0f * sqrtf(lmax)), MIN_RADIUS, MAX_RADIUS);

    // screen_space_center (243-257)
    s.px = k.fx * tx * inv_tz + k.cx;
    s.py = k.fy * ty * inv_tz + k.cy;
    s.conic[0] = conic0;
    s.conic[1] = conic1;
    s.conic[2] = conic2;
    // is_frustum_culled (261-274)
    const bool culled = (s.px + radius) < 0.0f || (s.px - radius) >= k.vw ||
                        (s.py + radius) < 0.0f || (s.py - radius) >= k.vh;
    s.radius = culled ? 0 : sat_i32(radius);
    return s;
}

struct Bounds {
    uint32_t tx0, tx1, ty0, ty1;
};

// tiles/mod.rs:97-130 (compute_gaussian_tile_bounds_xy with extent = radius)
inline bool tile_bounds(float cx, float cy, float r, uint32_t tiles_x, uint32_t tiles_y,
                        uint32_t W, uint32_t H, Bounds* out) {
    if (r <= 0.0f || tiles_x == 0 || tiles_y == 0) return false;
    const float min_x = cx - r, max_x = cx + r, min_y = cy - r, max_y = cy + r;
    if (max_x < 0.0f || min_x >= (float)W || max_y < 0.0f || min_y >= (float)H) return false;
    auto cl = [](float v, uint32_t n) {
        int32_t t = sat_i32(floorf(v / 16.0f));
        const int32_t hi = (int32_t)n - 1;
        if (t < 0) t = 0;
        if (t > hi) t = hi;
        return (uint32_t)t;
    };
    out->tx0 = cl(min_x, tiles_x);
    out->tx1 = cl(max_x, tiles_x);
    out->ty0 = cl(min_y, tiles_y);
    out->ty1 = cl(max_y, tiles_y);
    return !(out->tx1 < out->tx0 || out->ty1 < out->ty0);
}

struct Entry {
    uint32_t tile, depth_bits, id;
};

inline bool entry_less(const Entry& a, const Entry& b) {
    if (a.tile != b.tile) return a.tile < b.tile;
    if (a.depth_bits != b.depth_bits) return a.depth_bits < b.depth_bits;
    return a.id < b.id;
}

inline uint32_t f2u(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

// sh.rs:5-137 (evaluate_sh; 48 coefficients per splat, rest at 3 / 18 / 33)
void evaluate_sh(const float* c, V3 d, uint32_t degree, float rgb[3]) {
    const float C0 = 0.28209479177387814f, C1 = 0.4886025119029199f;
    const float C2_0 = 1.0925484305920792f, C2_1 = -1.0925484305920792f, C2_2 = 0.31539156525252005f,
                C2_3 = -1.0925484305920792f, C2_4 = 0.5462742152960396f;
    const float C3_0 = -0.5900435899266435f, C3_1 = 2.890611442640554f, C3_2 = -0.4570457994644658f,
                C3_3 = 0.3731763325901154f, C3_4 = -0.4570457994644658f, C3_5 = 1.445305721320277f,
                C3_6 = -0.5900435899266435f;
    float r = C0 * c[0], g = C0 * c[1], b = C0 * c[2];
    if (degree > 0) {
        const float x = d.x, y = d.y, z = d.z;
        const float y1m1 = -C1 * y, y10 = C1 * z, y1p1 = -C1 * x;
        r += (y1m1 * c[3] + y10 * c[4]) + y1p1 * c[5];
        g += (y1m1 * c[18] + y10 * c[19]) + y1p1 * c[20];
        b += (y1m1 * c[33] + y10 * c[34]) + y1p1 * c[35];
        if (degree >= 2) {
            const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
            const float y2m2 = C2_0 * xy, y2m1 = C2_1 * yz, y20 = C2_2 * ((2.0f * zz - xx) - yy);
            const float y2p1 = C2_3 * xz, y2p2 = C2_4 * (xx - yy);
            r += (((y2m2 * c[6] + y2m1 * c[7]) + y20 * c[8]) + y2p1 * c[9]) + y2p2 * c[10];
            g += (((y2m2 * c[21] + y2m1 * c[22]) + y20 * c[23]) + y2p1 * c[24]) + y2p2 * c[25];
            b += (((y2m2 * c[36] + y2m1 * c[37]) + y20 * c[38]) + y2p1 * c[39]) + y2p2 * c[40];
            if (degree >= 3) {
                const float y3m3 = C3_0 * y * (3.0f * xx - yy);
                const float y3m2 = C3_1 * xy * z;
                const float y3m1 = C3_2 * y * ((4.0f * zz - xx) - yy);
                const float y30 = C3_3 * z * ((2.0f * zz - 3.0f * xx) - 3.0f * yy);
                const float y3p1 = C3_4 * x * ((4.0f * zz - xx) - yy);
                const float y3p2 = C3_5 * z * (xx - yy);
                const float y3p3 = C3_6 * x * (xx - 3.0f * yy);
                r += (((((y3m3 * c[11] + y3m2 * c[12]) + y3m1 * c[13]) + y30 * c[14]) + y3p1 * c[15]) +
                      y3p2 * c[16]) + y3p3 * c[17];
                g += (((((y3m3 * c[26] + y3m2 * c[27]) + y3m1 * c[28]) + y30 * c[29]) + y3p1 * c[30]) +
                      y3p2 * c[31]) + y3p3 * c[32];
                b += (((((y3m3 * c[41] + y3m2 * c[42]) + y3m1 * c[43]) + y30 * c[44]) + y3p1 * c[45]) +
                      y3p2 * c[46]) + y3p3 * c[47];
            }
        }
    }
    rgb[0] = rclamp(r + 0.5f, 0.0f, 1.0f);
    rgb[1] = rclamp(g + 0.5f, 0.0f, 1.0f);
    rgb[2] = rclamp(b + 0.5f, 0.0f, 1.0f);
}

struct ThreadScope {
    int saved;
    explicit ThreadScope(int n) : saved(omp_get_max_threads()) { omp_set_num_threads(n < 1 ? 1 : n); }
    ~ThreadScope() { omp_set_num_threads(saved); }
};

void project_all(const float* pos, const float* cov, uint64_t n, const Cam& k, float* pts, float* depths,
                 int32_t* radii, float* conics) {
    const int64_t chunks = (int64_t)((n + PROJECT_CHUNK - 1) / PROJECT_CHUNK);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t ch = 0; ch < chunks; ++ch) {
        const uint64_t i0 = (uint64_t)ch * PROJECT_CHUNK;
        const uint64_t i1 = std::min<uint64_t>(n, i0 + PROJECT_CHUNK);
        for (uint64_t i = i0; i < i1; ++i) {
            const V3 p{pos[i * 3], pos[i * 3 + 1], pos[i * 3 + 2]};
            const Splat s = project_ewa_splat(p, cov + i * 6, k);
            pts[i * 2] = s.px;
            pts[i * 2 + 1] = s.py;
            depths[i] = s.depth;
            radii[i] = s.radius;
            conics[i * 3] = s.conic[0];
            conics[i * 3 + 1] = s.conic[1];
            conics[i * 3 + 2] = s.conic[2];
        }
    }
}

// cpu_reference.rs:66-92 (counts), 94-111 (serial prefix), 168-205 (scatter
// by binary search), 207-209 (sort), 211-226 (ranges).  Returns the entry
// count; entries / ranges are written when `entries` is non-null.
uint64_t bin_all(const float* pts, const float* depths, const int32_t* radii, uint64_t n, uint32_t W,
                 uint32_t H, std::vector<Entry>* entries, uint32_t* ranges, double* ms_sort) {
    const uint32_t tiles_x = (W + TILE - 1) / TILE, tiles_y = (H + TILE - 1) / TILE;
    const uint64_t T = (uint64_t)tiles_x * tiles_y;
    std::vector<uint32_t> counts(n), offsets(n);
#pragma omp parallel for schedule(static, 4096)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        Bounds b;
        const int32_t r = radii[i];
        counts[i] = (r > 0 && tile_bounds(pts[i * 2], pts[i * 2 + 1], (float)r, tiles_x, tiles_y, W, H, &b))
                        ? (b.tx1 - b.tx0 + 1) * (b.ty1 - b.ty0 + 1)
                        : 0u;
    }
    uint64_t running = 0;
    for (uint64_t i = 0; i < n; ++i) {
        offsets[i] = (uint32_t)running;
        running += counts[i];
    }
    const uint64_t total = running;
    if (!entries) return total;
    entries->assign(total, Entry{0, 0, 0});
    Entry* e = entries->data();
#pragma omp parallel for schedule(static, 8192)
    for (int64_t k = 0; k < (int64_t)total; ++k) {
        const uint32_t* ub = std::upper_bound(offsets.data(), offsets.data() + n, (uint32_t)k);
        const uint64_t owner = (uint64_t)(ub - offsets.data());
        if (owner == 0) continue;
        const uint64_t g = owner - 1;
        if (counts[g] == 0) continue;
        Bounds b;
        if (!tile_bounds(pts[g * 2], pts[g * 2 + 1], (float)std::max(radii[g], 0), tiles_x, tiles_y, W, H, &b))
            continue;
        const uint32_t ord = (uint32_t)(k - offsets[g]);
        const uint32_t width = (b.tx1 - b.tx0) + 1;
        const uint32_t ty = b.ty0 + ord / width, tx = b.tx0 + ord % width;
        if (ty > b.ty1 || tx > b.tx1) continue;
        e[k] = Entry{ty * tiles_x + tx, f2u(depths[g]), (uint32_t)g};
    }
    const auto t0 = std::chrono::steady_clock::now();
    if (omp_get_max_threads() > 1)
        __gnu_parallel::sort(e, e + total, entry_less);
    else
        std::sort(e, e + total, entry_less);
    if (ms_sort)
        *ms_sort = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
#pragma omp parallel for schedule(static, 256)
    for (int64_t t = 0; t < (int64_t)T; ++t) {
        const Entry* end = e + total;
        const Entry* lo = std::partition_point((const Entry*)e, end, [t](const Entry& x) { return x.tile < (uint32_t)t; });
        const Entry* hi = std::partition_point(lo, end, [t](const Entry& x) { return x.tile <= (uint32_t)t; });
        ranges[t * 2] = lo < hi ? (uint32_t)(lo - e) : 0u;
        ranges[t * 2 + 1] = lo < hi ? (uint32_t)(hi - e) : 0u;
    }
    return total;
}

// cpu_reference.rs:120-240 (rasterize_scene_cpu) + 53-70 (from_pixels)
void rasterize_all(const float* pts, const float* conics, const float* opa, const float* sh, const float* pos,
                   const Entry* e, const uint32_t* ranges, uint32_t W, uint32_t H, const float* bg,
                   const float* cam_pos, uint32_t sh_degree, uint8_t* rgba8, float* rgbaf) {
    const uint32_t tiles_x = (W + TILE - 1) / TILE;
    const uint32_t deg = sh_degree < 3 ? sh_degree : 3;
    const int64_t npix = (int64_t)W * H;
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t pi = 0; pi < npix; ++pi) {
        const uint32_t px = (uint32_t)(pi % W), py = (uint32_t)(pi / W);
        const uint32_t tile = (py / 16) * tiles_x + (px / 16);
        const uint32_t start = ranges[tile * 2], end = ranges[tile * 2 + 1];
        float out[4];
        if (start >= end) {
            out[0] = bg[0];
            out[1] = bg[1];
            out[2] = bg[2];
            out[3] = bg[3];
        } else {
            const float pcx = (float)px + 0.5f, pcy = (float)py + 0.5f;
            float cr = 0.0f, cg = 0.0f, cb = 0.0f, t = 1.0f;
            for (uint32_t k = start; k < end; ++k) {
                if (t < 0.001f) {
                    t = 0.0f;
                    break;
                }
                const uint32_t g = e[k].id;
                const float dx = pcx - pts[g * 2], dy = pcy - pts[g * 2 + 1];
                const float a = conics[g * 3], b = conics[g * 3 + 1], c = conics[g * 3 + 2];
                float q = -0.5f * ((a * dx * dx + 2.0f * b * dx * dy) + c * dy * dy);
                if (q > 0.0f) q = 0.0f;
                if (q < -4.0f) continue;
                const float alpha = rmin(opa[g] * expf(q), 0.99f);
                if (alpha < 1.0f / 255.0f) continue;
                const V3 to{pos[g * 3] - cam_pos[0], pos[g * 3 + 1] - cam_pos[1], pos[g * 3 + 2] - cam_pos[2]};
                const float len2 = (to.x * to.x + to.y * to.y) + to.z * to.z;
                V3 dir{0.0f, 0.0f, 1.0f};
                if (len2 > 1e-8f) {
                    // glam Vec3::normalize: self * (1 / sqrt(dot(self, self)))
                    const float rl = 1.0f / sqrtf(len2);
                    dir = V3{to.x * rl, to.y * rl, to.z * rl};
                }
                float rgb[3];
                evaluate_sh(sh + (uint64_t)g * 48, dir, deg, rgb);
                const float w = alpha * t;
                cr += w * rgb[0];
                cg += w * rgb[1];
                cb += w * rgb[2];
                t *= 1.0f - alpha;
            }
            out[0] = rclamp(cr + t * bg[0], 0.0f, 1.0f);
            out[1] = rclamp(cg + t * bg[1], 0.0f, 1.0f);
            out[2] = rclamp(cb + t * bg[2], 0.0f, 1.0f);
            out[3] = rclamp(1.0f - t, 0.0f, 1.0f);
        }
        for (int ch = 0; ch < 4; ++ch) {
            if (rgbaf) rgbaf[pi * 4 + ch] = out[ch];
            rgba8[pi * 4 + ch] = (uint8_t)roundf(rclamp(out[ch], 0.0f, 1.0f) * 255.0f);
        }
    }
}

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

// synthetic.rs:6-42
O1_API void o1_fibonacci_cluster(uint32_t count, float radius, const float* center, float* pos, float* cov,
                                 float* opa, float* sh) {
    const float golden = (1.0f + sqrtf(5.0f)) / 2.0f;
    const float SH_C0 = 0.28209479f;
    const float PI = 3.14159265358979323846f;
    for (uint32_t i = 0; i < count; ++i) {
        const float t = ((float)i + 0.5f) / (float)count;
        const float r = radius * cbrtf(t);
        const float theta = acosf(rclamp(1.0f - 2.0f * t, -1.0f, 1.0f));
        const float phi = 2.0f * PI * ((float)i / golden);
        pos[i * 3] = center[0] + r * sinf(theta) * cosf(phi);
        pos[i * 3 + 1] = center[1] + r * sinf(theta) * sinf(phi);
        pos[i * 3 + 2] = center[2] + r * cosf(theta);
        const float cv[6] = {0.0004f, 0.0f, 0.0f, 0.0004f, 0.0f, 0.0004f};
        memcpy(cov + (uint64_t)i * 6, cv, sizeof(cv));
        opa[i] = 0.88f;
        float* s = sh + (uint64_t)i * 48;
        memset(s, 0, 48 * sizeof(float));
        const float cr = 0.5f + 0.5f * sinf((float)i * 0.02f);
        const float cg = 0.5f + 0.5f * sinf((float)i * 0.02f + 2.094f);
        const float cb = 0.5f + 0.5f * sinf((float)i * 0.02f + 4.188f);
        s[0] = (cr - 0.5f) / SH_C0;
        s[1] = (cg - 0.5f) / SH_C0;
        s[2] = (cb - 0.5f) / SH_C0;
    }
}

// project_scene_cpu (render/mod.rs:143-212).  cam = CameraUniforms as 24 floats.
O1_API void o1_project(const float* pos, const float* cov, uint64_t n, const float* cam, float* pts,
                       float* depths, int32_t* radii, float* conics, int threads) {
    ThreadScope ts(threads);
    project_all(pos, cov, n, load_cam(cam), pts, depths, radii, conics);
}

// bin_and_sort_gaussians_cpu (tiles/cpu_reference.rs:248-302).  Returns the
// entry count; writes entries (tile, depth bits, id) and ranges (start, end)
// per tile only when the count fits `cap`.
O1_API uint64_t o1_bin(const float* pts, const float* depths, const int32_t* radii, uint64_t n, uint32_t W,
                       uint32_t H, uint32_t* entries, uint64_t cap, uint32_t* ranges, int threads) {
    ThreadScope ts(threads);
    const uint64_t total = bin_all(pts, depths, radii, n, W, H, nullptr, nullptr, nullptr);
    if (total > cap || !entries) return total;
    std::vector<Entry> e;
    bin_all(pts, depths, radii, n, W, H, &e, ranges, nullptr);
    memcpy(entries, e.data(), total * sizeof(Entry));
    return total;
}

// rasterize_scene_cpu (rasterizer/cpu_reference.rs:120-240); entries as o1_bin
// writes them.  rgbaf (optional) receives the unquantised pixels.
O1_API void o1_rasterize(const float* pts, const float* conics, const float* opa, const float* sh,
                         const float* pos, const uint32_t* entries, const uint32_t* ranges, uint32_t W,
                         uint32_t H, const float* bg, const float* cam_pos, uint32_t sh_degree, uint8_t* rgba8,
                         float* rgbaf, int threads) {
    ThreadScope ts(threads);
    rasterize_all(pts, conics, opa, sh, pos, reinterpret_cast<const Entry*>(entries), ranges, W, H, bg, cam_pos,
                  sh_degree, rgba8, rgbaf);
}

// render_scene_cpu (render/mod.rs:288-305) with per-stage wall times:
// stage_ms = {projection, binning (counts + prefix + scatter + sort + ranges),
// sort alone, compositor}; stats = {entries, surviving splats}.
O1_API void o1_render(const float* pos, const float* cov, const float* opa, const float* sh, uint64_t n,
                      const float* cam, uint32_t W, uint32_t H, const float* bg, const float* cam_pos,
                      uint32_t sh_degree, uint8_t* rgba8, double* stage_ms, uint64_t* stats, int threads) {
    ThreadScope ts(threads);
    const Cam k = load_cam(cam);
    const uint64_t T = (uint64_t)((W + TILE - 1) / TILE) * ((H + TILE - 1) / TILE);
    auto t0 = std::chrono::steady_clock::now();
    std::vector<float> pts(n * 2), depths(n), conics(n * 3);
    std::vector<int32_t> radii(n);
    project_all(pos, cov, n, k, pts.data(), depths.data(), radii.data(), conics.data());
    const double ms_proj = ms_since(t0);
    t0 = std::chrono::steady_clock::now();
    std::vector<Entry> e;
    std::vector<uint32_t> ranges(T * 2);
    double ms_sort = 0.0;
    const uint64_t total = bin_all(pts.data(), depths.data(), radii.data(), n, W, H, &e, ranges.data(), &ms_sort);
    const double ms_bin = ms_since(t0);
    t0 = std::chrono::steady_clock::now();
    rasterize_all(pts.data(), conics.data(), opa, sh, pos, e.data(), ranges.data(), W, H, bg, cam_pos, sh_degree,
                  rgba8, nullptr);
    const double ms_raster = ms_since(t0);
    if (stage_ms) {
        stage_ms[0] = ms_proj;
        stage_ms[1] = ms_bin;
        stage_ms[2] = ms_sort;
        stage_ms[3] = ms_raster;
    }
    if (stats) {
        uint64_t surv = 0;
        for (uint64_t i = 0; i < n; ++i) surv += radii[i] > 0;
        stats[0] = total;
        stats[1] = surv;
    }
}

// RasterizedImage::pixel_hash (rasterizer/cpu_reference.rs:105-112): FNV-1a.
O1_API uint64_t o1_pixel_hash(const uint8_t* rgba8, uint64_t len) {
    uint64_t h = 0xcbf29ce484222325ull;
    for (uint64_t i = 0; i < len; ++i) {
        h ^= rgba8[i];
        h *= 0x00000100000001b3ull;
    }
    return h;
}

// evaluate_sh (rasterizer/sh.rs:29-137) on one splat's 48 coefficients.
O1_API void o1_evaluate_sh(const float* coeffs48, const float* dir, uint32_t degree, float* rgb) {
    evaluate_sh(coeffs48, V3{dir[0], dir[1], dir[2]}, degree, rgb);
}
