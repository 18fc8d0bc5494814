// spectral.cpp -- host spectral helpers of scene ingestion.
//
// The colour-matching functions are the Wyman-Sloan-Shirley 2013 analytic fit
// (code/Includes/CIE.h:27-53), the sRGB matrices are IEC 61966-2-1 (CIE.h:97-123),
// and the RGB -> spectrum fit is Jakob & Hanika 2019's sigmoid of a quadratic,
// solved by Newton-Raphson with a chromaticity homotopy from grey and a spike
// guard (code/RGBToSpectrum.cpp:27-293).  Scene files are fitted once at load
// and the fits travel to the device; they must equal the reference's bit for
// bit, so operation order and libm calls (expf, sqrtf) follow the reference.
#include <algorithm>
#include <cmath>
#include <fstream>
#include <sstream>

#include "scene.h"

namespace pcr {

namespace {

constexpr int kN = 61;
constexpr float kLo = 400.f, kHi = 700.f;
constexpr float kDl = (kHi - kLo) / (kN - 1);
inline float grid(int i) { return kLo + i * kDl; }

inline float gauss(float l, float mu, float a, float b) {
    float t = (l < mu) ? (l - mu) * a : (l - mu) * b;
    return std::exp(-0.5f * t * t);
}
inline float cmf_x(float l) {
    return 0.362f * gauss(l, 442.0f, 0.0624f, 0.0374f) + 1.056f * gauss(l, 599.8f, 0.0264f, 0.0323f) -
           0.065f * gauss(l, 501.1f, 0.0490f, 0.0382f);
}
inline float cmf_y(float l) {
    return 0.821f * gauss(l, 568.8f, 0.0213f, 0.0247f) + 0.286f * gauss(l, 530.9f, 0.0613f, 0.0322f);
}
inline float cmf_z(float l) {
    return 1.217f * gauss(l, 437.0f, 0.0845f, 0.0278f) + 0.681f * gauss(l, 459.0f, 0.0385f, 0.0725f);
}

inline float sig(float x) { return 0.5f + x / (2.f * std::sqrt(1.f + x * x)); }
inline float dsig(float x) {
    float t = 1.f + x * x;
    return 0.5f / (t * std::sqrt(t));
}
inline float sig_inv(float y) {
    y = std::clamp(y, 1e-4f, 1.f - 1e-4f);
    float u = 2.f * y - 1.f;
    return u / std::sqrt(1.f - u * u);
}
inline float norm_lambda(int i) {
    constexpr float mid = 0.5f * (kLo + kHi), half = 0.5f * (kHi - kLo);
    return (grid(i) - mid) / half;
}

// Integrated XYZ of sigmoid(c0 + c1 L + c2 L^2) on the grid (and its Jacobian).
void integrate(const float c[3], float xyz[3], float J[3][3]) {
    float X = 0.f, Y = 0.f, Z = 0.f;
    float j[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int i = 0; i < kN; i++) {
        float l = grid(i);
        float L = norm_lambda(i);
        float p = c[0] + L * (c[1] + L * c[2]);
        float s = sig(p);
        float xb = cmf_x(l), yb = cmf_y(l), zb = cmf_z(l);
        X += s * xb;
        Y += s * yb;
        Z += s * zb;
        if (J) {
            float ds = dsig(p), L2 = L * L;
            j[0][0] += ds * xb;
            j[0][1] += ds * L * xb;
            j[0][2] += ds * L2 * xb;
            j[1][0] += ds * yb;
            j[1][1] += ds * L * yb;
            j[1][2] += ds * L2 * yb;
            j[2][0] += ds * zb;
            j[2][1] += ds * L * zb;
            j[2][2] += ds * L2 * zb;
        }
    }
    xyz[0] = X * kDl;
    xyz[1] = Y * kDl;
    xyz[2] = Z * kDl;
    if (J)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) J[a][b] = j[a][b] * kDl;
}

// Cramer's rule (RGBToSpectrum.cpp:27-53)
bool cramer(const float J[3][3], const float b[3], float x[3]) {
    float det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    if (std::abs(det) < 1e-12f) return false;
    float inv = 1.f / det;
    float d0 = b[0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (b[1] * J[2][2] - J[1][2] * b[2]) +
               J[0][2] * (b[1] * J[2][1] - J[1][1] * b[2]);
    float d1 = J[0][0] * (b[1] * J[2][2] - J[1][2] * b[2]) - b[0] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
               J[0][2] * (J[1][0] * b[2] - b[1] * J[2][0]);
    float d2 = J[0][0] * (J[1][1] * b[2] - b[1] * J[2][1]) - J[0][1] * (J[1][0] * b[2] - b[1] * J[2][0]) +
               b[0] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    x[0] = d0 * inv;
    x[1] = d1 * inv;
    x[2] = d2 * inv;
    return true;
}

float residual_norm(const float a[3], const Vec3 &t) {
    float r0 = a[0] - t.x, r1 = a[1] - t.y, r2 = a[2] - t.z;
    return std::sqrt(r0 * r0 + r1 * r1 + r2 * r2);
}

// Damped Newton (at most 15 steps, 4 halvings each; RGBToSpectrum.cpp:59-158).
void newton(const Vec3 &target, float c[3]) {
    for (int it = 0; it < 15; it++) {
        float xyz[3], J[3][3];
        integrate(c, xyz, J);
        float r[3] = {xyz[0] - target.x, xyz[1] - target.y, xyz[2] - target.z};
        float rn = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        if (rn < 1e-6f) return;
        float delta[3];
        if (!cramer(J, r, delta)) return;
        float prev[3] = {c[0], c[1], c[2]};
        float step = 1.f;
        bool improved = false;
        for (int k = 0; k < 4; k++) {
            for (int q = 0; q < 3; q++) c[q] = prev[q] - step * delta[q];
            float t[3];
            integrate(c, t, nullptr);
            if (residual_norm(t, target) < rn) {
                improved = true;
                break;
            }
            step *= 0.5f;
        }
        if (!improved) {
            for (int q = 0; q < 3; q++) c[q] = prev[q];
            return;
        }
    }
}

// max / median of the 61 samples (RGBToSpectrum.cpp:167-183)
float spike_ratio(const float c[3]) {
    float v[kN];
    float mx = 0.f;
    for (int i = 0; i < kN; i++) {
        float L = norm_lambda(i);
        float p = c[0] + L * (c[1] + L * c[2]);
        v[i] = 0.5f + p / (2.f * std::sqrt(1.f + p * p));
        mx = std::max(mx, v[i]);
    }
    std::nth_element(v, v + kN / 2, v + kN);
    float med = v[kN / 2];
    if (med < 1e-8f) return mx < 1e-8f ? 1.f : 1e30f;
    return mx / med;
}

}  // namespace

float ybar_integral() {
    static const float v = [] {
        float s = 0.f;
        for (int i = 0; i < kN; i++) s += cmf_y(grid(i));
        return s * kDl;
    }();
    return v;
}

Vec3 spectrum_to_xyz(const std::array<float, 61> &s) {
    float X = 0.f, Y = 0.f, Z = 0.f;
    for (int i = 0; i < kN; i++) {
        float l = grid(i);
        X += s[i] * cmf_x(l);
        Y += s[i] * cmf_y(l);
        Z += s[i] * cmf_z(l);
    }
    float k = kDl / ybar_integral();
    return Vec3{X * k, Y * k, Z * k};
}

Vec3 xyz_to_linear_srgb(Vec3 c) {
    return Vec3{3.2404542f * c.x - 1.5371385f * c.y - 0.4985314f * c.z,
                -0.9692660f * c.x + 1.8760108f * c.y + 0.0415560f * c.z,
                0.0556434f * c.x - 0.2040259f * c.y + 1.0572252f * c.z};
}
Vec3 linear_srgb_to_xyz(Vec3 c) {
    return Vec3{0.4124564f * c.x + 0.3575761f * c.y + 0.1804375f * c.z,
                0.2126729f * c.x + 0.7151522f * c.y + 0.0721750f * c.z,
                0.0193339f * c.x + 0.1191920f * c.y + 0.9503041f * c.z};
}

SigmoidFit fit_sigmoid(Vec3 rgb) {
    float r = std::clamp(rgb.x, 0.f, 1.f), g = std::clamp(rgb.y, 0.f, 1.f), b = std::clamp(rgb.z, 0.f, 1.f);
    SigmoidFit f;
    if (r <= 0.f && g <= 0.f && b <= 0.f) return f;
    if (r >= 1.f && g >= 1.f && b >= 1.f) {
        f.scale = ybar_integral();
        f.c0 = sig_inv(1.f / f.scale);
        return f;
    }
    Vec3 t = linear_srgb_to_xyz(Vec3{r, g, b});
    // homotopy: 32 steps from grey of equal luminance to the target
    const float Xn = 0.95046f, Zn = 1.08906f;
    Vec3 grey{t.y * Xn, t.y, t.y * Zn};
    float c[3] = {sig_inv(t.y / ybar_integral()), 0.f, 0.f};
    float keep[3] = {c[0], c[1], c[2]};
    for (int k = 1; k <= 32; k++) {
        float w = (float)k / (float)32;
        Vec3 cur{(1.f - w) * grey.x + w * t.x, (1.f - w) * grey.y + w * t.y, (1.f - w) * grey.z + w * t.z};
        float n[3] = {c[0], c[1], c[2]};
        newton(cur, n);
        if (spike_ratio(n) > 200.f) break;
        for (int q = 0; q < 3; q++) c[q] = keep[q] = n[q];
    }
    f.c0 = keep[0];
    f.c1 = keep[1];
    f.c2 = keep[2];
    f.scale = ybar_integral();
    return f;
}

// Material::populateSpectra (Material.h:125-139)
void populate_spectra(Material &m) {
    m.fitAlbedo = fit_sigmoid(m.albedo);
    if (m.isEmissive()) {
        float e = std::max({m.emissive.x, m.emissive.y, m.emissive.z});
        m.fitEmissive = fit_sigmoid(Vec3{m.emissive.x / e, m.emissive.y / e, m.emissive.z / e});
        m.fitEmissive.scale *= e;
    } else {
        m.fitEmissive = SigmoidFit{};
    }
}

// SpdLoader::loadSpdFile (SpdLoader.cpp:38-80): "lambda value" lines, '#'
// comments, ascending wavelengths, endpoint-clamped linear resampling.
bool load_spd_file(const std::string &path, std::array<float, 61> &out, std::string *err) {
    std::ifstream in(path);
    if (!in) {
        if (err) *err = "cannot open file: " + path;
        return false;
    }
    std::vector<float> lam, val;
    std::string line;
    int no = 0;
    while (std::getline(in, line)) {
        no++;
        size_t s = line.find_first_not_of(" \t");
        if (s == std::string::npos || line[s] == '#') continue;
        std::istringstream is(line.substr(s));
        float a, b;
        if (!(is >> a >> b)) {
            if (err) *err = path + ":" + std::to_string(no) + ": expected 'wavelength value'";
            return false;
        }
        lam.push_back(a);
        val.push_back(b);
    }
    if (lam.empty()) {
        if (err) *err = "no data samples in " + path;
        return false;
    }
    for (size_t i = 1; i < lam.size(); i++)
        if (lam[i] <= lam[i - 1]) {
            if (err) *err = path + ": wavelengths must be strictly ascending";
            return false;
        }
    for (int i = 0; i < kN; i++) {
        float l = grid(i);
        if (l <= lam.front()) out[i] = val.front();
        else if (l >= lam.back()) out[i] = val.back();
        else {
            size_t hi = (size_t)(std::upper_bound(lam.begin(), lam.end(), l) - lam.begin());
            size_t lo = hi - 1;
            float t = (l - lam[lo]) / (lam[hi] - lam[lo]);
            out[i] = val[lo] + t * (val[hi] - val[lo]);
        }
    }
    return true;
}

}  // namespace pcr
