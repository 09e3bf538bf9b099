// vbt_synth.cpp -- synthetic fields + VBTPACK4 encoder (asset tooling).
//
// Field shapes follow the benchmark configurations: advected Gaussian smoke
// blobs, a fire triple (density / narrow-band flames / temperature), a falling
// sphere level set merged by smooth-min, and a smooth sum-of-sines scientific
// field with sparse spikes.  Randomness is a seeded splitmix64 stream with
// explicit mappings, so assets are identical on every machine.
//
// The encoder writes the stream layout the decoder reads:
//   render leaf  = u32 header | u16 coarseKeys u16 fineKeys | coarse stream
//                  [| fine6 stream | 80-byte shell occupancy + shell stream]
//   stream       = desc[C]{u16 keyStart,u8 keyCount,u8 localMaxBinKeys}
//                  | u8 bins[C][B] (pad 4) | keyframes (pad 4)
//   scientific   = u32 header | 64 x {f16 DC, i8 AC[keep-1]} | residual block
// Keyframes come from an error-bounded greedy split of each control's time
// series (the role the reference's DP keyframe detector plays), bins from the
// "last key before the bin start" rule of the reference packer.
#include "vbt_synth.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;

struct Rng {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double u01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uni(double a, double b) { return a + (b - a) * u01(); }
};

inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t hash4(uint64_t seed, int64_t a, int64_t b, int64_t c, int64_t d)
{
    uint64_t h = mix64(seed ^ 0x51ED270B27A3C1F3ull);
    h = mix64(h ^ (uint64_t)a);
    h = mix64(h ^ ((uint64_t)b << 1));
    h = mix64(h ^ ((uint64_t)c << 2));
    h = mix64(h ^ ((uint64_t)d << 3));
    return h;
}

inline double hash01(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

// round-to-nearest-even float -> IEEE half bits
uint16_t to_half(float f)
{
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t absx = x & 0x7FFFFFFFu;
    if (absx >= 0x7F800000u) return (uint16_t)(sign | 0x7C00u | (absx > 0x7F800000u ? 0x200u : 0u));
    if (absx >= 0x477FF000u) return (uint16_t)(sign | 0x7C00u); // rounds to inf
    if (absx < 0x38800000u) {                                     // subnormal half
        if (absx < 0x33000000u) return (uint16_t)sign;
        const uint32_t e = absx >> 23;
        const uint32_t m = (absx & 0x7FFFFFu) | 0x800000u;
        const uint32_t shift = 126u - e; // 14..24
        uint32_t v = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1u);
        if (rem > half || (rem == half && (v & 1u))) ++v;
        return (uint16_t)(sign | v);
    }
    uint32_t v = ((absx - 0x38000000u) >> 13);
    const uint32_t rem = absx & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (v & 1u))) ++v;
    return (uint16_t)(sign | v);
}

float from_half(uint16_t h)
{
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1Fu, m = h & 0x3FFu, out;
    if (e == 0) {
        if (m == 0) out = sign;
        else {
            e = 1;
            while (!(m & 0x400u)) { m <<= 1; --e; }
            out = sign | ((e + 112u) << 23) | ((m & 0x3FFu) << 13);
        }
    } else if (e == 31) out = sign | 0x7F800000u | (m << 13);
    else out = sign | ((e + 112u) << 23) | (m << 13);
    float f;
    std::memcpy(&f, &out, 4);
    return f;
}

struct Blob {
    double c0[3], v[3], a[3];
    double omega, phase, sigma, peak;
};

struct Sphere {
    double x0, y0, z0, vx, vy, r, g, floorZ;
};

struct Wave {
    double k[3], w, phase, amp;
};

struct Field {
    vbts_params p;
    std::vector<Blob> blobs;
    std::vector<Sphere> spheres;
    std::vector<Wave> waves;
    double floorLevel = 0.0;
    double amp = 1.0;
    std::vector<double> ctab; // per integer frame: blob / sphere centres (x, y, z)

    void center(const Blob& b, double t, double c[3]) const
    {
        for (int i = 0; i < 3; ++i) c[i] = b.c0[i] + b.v[i] * t + b.a[i] * std::sin(b.omega * t + b.phase + i);
    }

    void sphere_center(const Sphere& s, double t, double c[3]) const
    {
        c[0] = s.x0 + s.vx * t;
        c[1] = s.y0 + s.vy * t;
        c[2] = std::max(s.floorZ, s.z0 - 0.5 * s.g * t * t);
    }

    void fill_centres()
    {
        const size_t n = blobs.empty() ? spheres.size() : blobs.size();
        ctab.assign((size_t)p.frames * n * 3, 0.0);
        for (uint32_t t = 0; t < p.frames; ++t)
            for (size_t i = 0; i < n; ++i) {
                double* c = &ctab[((size_t)t * n + i) * 3];
                if (!blobs.empty()) center(blobs[i], (double)t, c);
                else sphere_center(spheres[i], (double)t, c);
            }
    }

    const double* centre(size_t i, double t, double tmp[3]) const
    {
        const size_t n = blobs.empty() ? spheres.size() : blobs.size();
        const double ti = std::floor(t);
        if (ti == t && t >= 0 && t < (double)p.frames && !ctab.empty()) return &ctab[((size_t)t * n + i) * 3];
        if (!blobs.empty()) center(blobs[i], t, tmp);
        else sphere_center(spheres[i], t, tmp);
        return tmp;
    }

    template <class Fn> void for_each_item(size_t n, const std::vector<int>* subset, Fn&& fn) const
    {
        if (subset) {
            for (int i : *subset) fn((size_t)i);
        } else {
            for (size_t i = 0; i < n; ++i) fn(i);
        }
    }

    // raw sum of truncated Gaussians; smooth != 0 widens sigma by 2 voxels (Gaussian convolution)
    double blobs_at(double x, double y, double z, double t, bool smooth, const std::vector<int>* sub = nullptr) const
    {
        double s = 0.0;
        for_each_item(blobs.size(), sub, [&](size_t i) {
            const Blob& b = blobs[i];
            double tmp[3];
            const double* c = centre(i, t, tmp);
            const double dx = x - c[0], dy = y - c[1], dz = z - c[2];
            const double r2 = dx * dx + dy * dy + dz * dz;
            double sg = b.sigma, pk = b.peak;
            if (smooth) {
                const double s2 = std::sqrt(sg * sg + 4.0);
                pk *= (sg / s2) * (sg / s2) * (sg / s2);
                sg = s2;
            }
            if (r2 >= p.support * p.support * sg * sg) return; // compact support
            s += pk * std::exp(-r2 / (2.0 * sg * sg));
        });
        return s;
    }

    static double band(double d) // narrow band of the density, peak 1 at d = 0.4
    {
        const double lo = 0.3, hi = 0.5, mid = 0.4, half = 0.1;
        if (d <= lo || d >= hi) return 0.0;
        return 1.0 - std::fabs(d - mid) / half;
    }

    double density(double x, double y, double z, double t, const std::vector<int>* sub = nullptr) const
    {
        double v = blobs_at(x, y, z, t, false, sub);
        if (v > 0.0 && p.noise > 0.0f) {
            const uint64_t h = hash4(p.seed, (int64_t)std::llround(x * 4.0), (int64_t)std::llround(y * 4.0),
                                     (int64_t)std::llround(z * 4.0), (int64_t)std::llround(t));
            v *= 1.0 + (double)p.noise * (2.0 * hash01(h) - 1.0);
        }
        return v < 1e-3 ? 0.0 : v;
    }

    double levelset(double x, double y, double z, double t, const std::vector<int>* sub = nullptr) const
    {
        double phi = z - floorLevel;
        const double k = 3.0;
        for_each_item(spheres.size(), sub, [&](size_t i) {
            double tmp[3];
            const double* c = centre(i, t, tmp);
            const double d = std::sqrt((x - c[0]) * (x - c[0]) + (y - c[1]) * (y - c[1]) + (z - c[2]) * (z - c[2])) -
                             spheres[i].r;
            const double hh = std::max(k - std::fabs(phi - d), 0.0) / k;
            phi = std::min(phi, d) - hh * hh * k * 0.25;
        });
        return phi;
    }

    double smooth_sci(double x, double y, double z, double t) const
    {
        double s = 0.0;
        for (const auto& w : waves) s += w.amp * std::sin(w.k[0] * x + w.k[1] * y + w.k[2] * z + w.w * t + w.phase);
        return s;
    }

    double spike(int64_t x, int64_t y, int64_t z, int64_t t) const
    {
        const uint64_t h = hash4(p.seed ^ 0xABCDEF12345ull, x, y, z, t);
        if (hash01(h) >= (double)p.spikeFraction) return 0.0;
        return ((h & 1u) ? 1.0 : -1.0) * (double)p.spikeGain * amp;
    }

    // value at a continuous position; spikes only at integer voxels (voxel == true)
    double eval(double x, double y, double z, double t, bool voxel, const std::vector<int>* sub = nullptr) const
    {
        switch (p.kind) {
        case VBTS_DENSITY: return density(x, y, z, t, sub);
        case VBTS_FLAMES: return band(density(x, y, z, t, sub));
        case VBTS_TEMPERATURE: return 300.0 + 1500.0 * band(blobs_at(x, y, z, t, true, sub));
        case VBTS_LEVELSET: return levelset(x, y, z, t, sub);
        default: {
            double v = smooth_sci(x, y, z, t);
            if (voxel) v += spike((int64_t)x, (int64_t)y, (int64_t)z, (int64_t)t);
            return v;
        }
        }
    }

    // can a leaf box [x0,x0+7]^3 be nonzero at any frame? (blob-support fields only)
    bool leaf_may_be_nonzero(int x0, int y0, int z0) const
    {
        if (p.kind != VBTS_DENSITY && p.kind != VBTS_FLAMES) return true;
        for (uint32_t t = 0; t < p.frames; ++t)
            for (size_t i = 0; i < blobs.size(); ++i) {
                const Blob& b = blobs[i];
                double tmp[3];
                const double* c = centre(i, (double)t, tmp);
                const double lo[3] = {(double)x0, (double)y0, (double)z0};
                double d2 = 0.0;
                for (int i = 0; i < 3; ++i) {
                    const double q = std::min(std::max(c[i], lo[i]), lo[i] + 7.0);
                    d2 += (c[i] - q) * (c[i] - q);
                }
                if (d2 < (double)p.support * p.support * b.sigma * b.sigma) return true;
            }
        return false;
    }
};

// Items (blobs / spheres) that can influence a leaf box [x0, x0+7]^3 at any
// frame.  Blobs: support sphere (2 sigma, widened for the smoothed field)
// reaches the box.  Spheres: the SDF term can enter the smooth-min within the
// box (1-Lipschitz terms, k = 3 blend width, 6.1 = box half diagonal + margin).
std::vector<int> leaf_subset(const Field& f, int x0, int y0, int z0)
{
    std::vector<int> out;
    const double lo[3] = {(double)x0, (double)y0, (double)z0};
    if (!f.blobs.empty()) {
        for (size_t i = 0; i < f.blobs.size(); ++i) {
            double sg = f.blobs[i].sigma;
            if (f.p.kind == VBTS_TEMPERATURE) sg = std::sqrt(sg * sg + 4.0);
            bool hit = false;
            for (uint32_t t = 0; t < f.p.frames && !hit; ++t) {
                double tmp[3];
                const double* c = f.centre(i, (double)t, tmp);
                double d2 = 0.0;
                for (int a = 0; a < 3; ++a) {
                    const double q = std::min(std::max(c[a], lo[a]), lo[a] + 7.0);
                    d2 += (c[a] - q) * (c[a] - q);
                }
                hit = d2 < (double)f.p.support * f.p.support * sg * sg;
            }
            if (hit) out.push_back((int)i);
        }
    } else if (!f.spheres.empty()) {
        const double cc[3] = {x0 + 3.5, y0 + 3.5, z0 + 3.5};
        std::vector<char> keep(f.spheres.size(), 0);
        for (uint32_t t = 0; t < f.p.frames; ++t) {
            double best = cc[2] - f.floorLevel;
            std::vector<double> d(f.spheres.size());
            for (size_t i = 0; i < f.spheres.size(); ++i) {
                double tmp[3];
                const double* c = f.centre(i, (double)t, tmp);
                d[i] = std::sqrt((cc[0] - c[0]) * (cc[0] - c[0]) + (cc[1] - c[1]) * (cc[1] - c[1]) +
                                 (cc[2] - c[2]) * (cc[2] - c[2])) - f.spheres[i].r;
                best = std::min(best, d[i]);
            }
            for (size_t i = 0; i < f.spheres.size(); ++i)
                if (d[i] < best + 2.0 * 6.1 + 3.0 + 1.0) keep[i] = 1;
        }
        for (size_t i = 0; i < keep.size(); ++i)
            if (keep[i]) out.push_back((int)i);
    }
    return out;
}

void build_field(Field& f)
{
    const vbts_params& p = f.p;
    Rng rng{p.seed};
    const double dims[3] = {(double)p.width, (double)p.height, (double)p.depth};
    const double F = std::max(1.0, (double)p.frames);
    if (p.kind <= VBTS_TEMPERATURE) {
        for (uint32_t i = 0; i < p.count; ++i) {
            Blob b{};
            for (int a = 0; a < 3; ++a) {
                b.c0[a] = rng.uni(0.15, 0.85) * dims[a];
                b.v[a] = rng.uni(-1.0, 1.0) * p.motion * 0.25 * dims[a] / F;
                b.a[a] = rng.uni(0.0, 1.0) * p.motion * 0.04 * dims[a];
            }
            b.omega = rng.uni(0.05, 0.25);
            b.phase = rng.uni(0.0, 6.283185307179586);
            b.sigma = rng.uni(p.sizeMin, p.sizeMax);
            b.peak = rng.uni(p.peakMin, p.peakMax);
            f.blobs.push_back(b);
        }
    } else if (p.kind == VBTS_LEVELSET) {
        f.floorLevel = 0.06 * dims[2];
        for (uint32_t i = 0; i < p.count; ++i) {
            Sphere s{};
            s.r = rng.uni(p.sizeMin, p.sizeMax);
            s.x0 = rng.uni(0.12, 0.88) * dims[0];
            s.y0 = rng.uni(0.12, 0.88) * dims[1];
            s.z0 = rng.uni(0.35, 0.92) * dims[2];
            s.vx = rng.uni(-1.0, 1.0) * 0.05 * dims[0] / F;
            s.vy = rng.uni(-1.0, 1.0) * 0.05 * dims[1] / F;
            s.g = 2.0 * 0.8 * dims[2] / (F * F) * rng.uni(0.5, 1.5);
            s.floorZ = f.floorLevel + s.r * 0.6;
            f.spheres.push_back(s);
        }
    } else {
        double sum = 0.0;
        for (uint32_t i = 0; i < p.count; ++i) {
            Wave w{};
            const double lambda = rng.uni(24.0, 96.0);
            double dir[3] = {rng.uni(-1, 1), rng.uni(-1, 1), rng.uni(-1, 1)};
            const double n = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]) + 1e-9;
            for (int a = 0; a < 3; ++a) w.k[a] = dir[a] / n * 6.283185307179586 / lambda;
            w.w = rng.uni(0.01, 0.08);
            w.phase = rng.uni(0.0, 6.283185307179586);
            w.amp = rng.uni(0.5, 1.0);
            sum += w.amp;
            f.waves.push_back(w);
        }
        for (auto& w : f.waves) w.amp /= std::max(1e-9, sum) / 2.0;
        f.amp = 1.0;
    }
    f.fill_centres();
}

// ---------------------------------------------------------------- encoder --
struct Stream {
    std::vector<uint8_t> desc, bins, keys; // keys: packed keyframe bytes
    uint32_t keyCount = 0;
};

// Keyframe detector of one series, as the reference's render encoder keys
// density and control series (detectDensityKeyFrames with its per-frame
// tolerance, render_temporal_kf_probe_main.cpp:200-275, 2080-2167): start from
// the two ends, split every span at the frame whose linear-interpolation miss
// is largest relative to that frame's tolerance until none exceeds it, and
// keep frames in the render cutoff band and on both sides of every crossing.
struct KeyProfile {
    float epsAbs, epsRel, gammaDelta, cutoff, band;
};

int cutoff_state(float v, const KeyProfile& kp)
{
    if (kp.cutoff <= 0.0f) return 1;
    if (v < kp.cutoff - kp.band) return 0;
    if (v > kp.cutoff + kp.band) return 2;
    return 1;
}

// Frames at which `v` is within `band` of 0 or changes sign, and their
// neighbours (addIsoSurfaceKeyFrames, render_temporal_kf_probe_main.cpp:2052-2078).
void mark_iso_keys(const float* v, int F, float band, std::vector<char>& on)
{
    band = std::max(0.0f, band);
    auto mark3 = [&](int t) {
        for (int u = t - 1; u <= t + 1; ++u)
            if (u >= 0 && u < F) on[u] = 1;
    };
    for (int t = 0; t < F; ++t) {
        if (std::fabs(v[t]) <= band) mark3(t);
        if (t > 0 && (v[t - 1] >= 0.0f) != (v[t] >= 0.0f)) {
            mark3(t - 1);
            mark3(t);
        }
    }
}

// iso: when set, the iso-surface keys of that series (band isoBand) join the
// series' own keys, as the reference adds them to level-set voxel and shell
// series
void fit_keys(const float* s, int F, const KeyProfile& kp, std::vector<int>& keys, const float* iso = nullptr,
              float isoBand = 0.0f)
{
    keys.clear();
    if (F <= 0) return;
    if (F == 1) {
        keys.push_back(0);
        return;
    }
    std::vector<char> on(F, 0);
    auto finish = [&]() {
        if (iso) mark_iso_keys(iso, F, isoBand, on);
        for (int t = 0; t < F; ++t)
            if (on[t]) keys.push_back(t);
    };
    double sum = 0.0, sum2 = 0.0;
    for (int t = 0; t < F; ++t) {
        sum += s[t];
        sum2 += (double)s[t] * s[t];
    }
    const float floorEps = std::max(1e-6f, kp.epsAbs * 0.1f);
    const double dmean = sum / F;
    if (std::max(0.0, sum2 / F - dmean * dmean) < (double)floorEps * floorEps) { // constant: one key
        on[0] = 1;
        finish();
        return;
    }
    const float mean = (float)dmean;
    const float sd = (float)std::sqrt(std::max(0.0, sum2 / F - (double)mean * mean));
    std::vector<float> steps(F - 1);
    for (int t = 1; t < F; ++t) steps[t - 1] = std::fabs(s[t] - s[t - 1]);
    std::sort(steps.begin(), steps.end());
    const float stepScale = std::max(1e-6f, steps[std::min(F - 2, (F - 1) * 9 / 10)]);
    std::vector<float> tol(F);
    for (int t = 0; t < F; ++t) {
        float e = kp.epsAbs + kp.epsRel * std::max(std::fabs(s[t]), sd);
        if (kp.gammaDelta > 0.0f && t > 0)
            e /= 1.0f + kp.gammaDelta * (std::fabs(s[t] - s[t - 1]) / stepScale);
        tol[t] = std::max(e, floorEps);
    }
    on[0] = on[F - 1] = 1;
    std::vector<std::pair<int, int>> stack{{0, F - 1}};
    while (!stack.empty()) {
        auto [a, b] = stack.back();
        stack.pop_back();
        if (b - a < 2) continue;
        double worst = 0.0;
        int wt = -1;
        const double v0 = s[a], v1 = s[b];
        for (int t = a + 1; t < b; ++t) {
            const double pred = v0 + (double)(t - a) / (double)(b - a) * (v1 - v0);
            const double e = std::fabs((double)s[t] - pred) / (double)tol[t];
            if (e > worst) {
                worst = e;
                wt = t;
            }
        }
        if (wt >= 0 && worst > 1.0) {
            on[wt] = 1;
            stack.push_back({a, wt});
            stack.push_back({wt, b});
        }
    }
    if (kp.cutoff > 0.0f && kp.band >= 0.0f) {
        auto mark = [&](int t) {
            if (t >= 0 && t < F) on[t] = 1;
        };
        for (int t = 0; t < F; ++t) {
            const int st = cutoff_state(s[t], kp);
            if (st == 1 || (t > 0 && cutoff_state(s[t - 1], kp) != st)) {
                mark(t - 1);
                mark(t);
                mark(t + 1);
            }
        }
    }
    finish();
}

// The series its keys interpolate (reconstructFromKeys,
// render_temporal_kf_probe_main.cpp:2169-2200): exact values at the keys,
// linear between them, held before the first and after the last.
void recon_linear(const float* s, const std::vector<int>& keys, int F, float* out)
{
    if (keys.empty()) {
        std::copy(s, s + F, out);
        return;
    }
    size_t seg = 0;
    for (int t = 0; t < F; ++t) {
        while (seg + 1 < keys.size() && t > keys[seg + 1]) ++seg;
        if (seg + 1 >= keys.size()) {
            out[t] = s[keys.back()];
            continue;
        }
        const int a = keys[seg], b = keys[seg + 1];
        out[t] = t <= a ? s[a] : (float)((double)s[a] + (double)(t - a) / (double)(b - a) * (double)(s[b] - s[a]));
    }
}

// bins: buildBinIndexFromKeys semantics (last key before each bin start)
void append_control(Stream& st, const float* series, std::vector<int>& keys, int F, int B, int codec)
{
    const uint32_t keyStart = st.keyCount;
    const size_t K = std::min<size_t>(keys.size(), 255);
    keys.resize(K);
    st.desc.push_back((uint8_t)(keyStart & 0xFF));
    st.desc.push_back((uint8_t)(keyStart >> 8));
    st.desc.push_back((uint8_t)K);
    std::vector<uint8_t> b(B, 0);
    for (int bin = 0; bin < B; ++bin) {
        const int start = (int)(((int64_t)bin * F) / std::max(1, B));
        uint8_t ls = (uint8_t)(K - 1);
        for (size_t i = 0; i < K; ++i)
            if (keys[i] >= start) { ls = (uint8_t)(i > 0 ? i - 1 : 0); break; }
        b[bin] = ls;
    }
    int localMax = 1;
    for (int bin = 0; bin < B; ++bin) {
        const int next = bin + 1 < B ? b[bin + 1] : (int)K;
        localMax = std::max(localMax, std::max(1, next - (int)b[bin]));
    }
    st.desc.push_back((uint8_t)std::min(localMax, 255));
    st.bins.insert(st.bins.end(), b.begin(), b.end());
    for (int kf : keys) {
        if (codec == 1) {
            uint8_t rec[8] = {(uint8_t)(kf & 0xFF), (uint8_t)(kf >> 8), 0, 0, 0, 0, 0, 0};
            std::memcpy(rec + 4, &series[kf], 4);
            st.keys.insert(st.keys.end(), rec, rec + 8);
        } else {
            const uint16_t h = to_half(series[kf]);
            st.keys.push_back((uint8_t)kf);
            st.keys.push_back((uint8_t)(h & 0xFF));
            st.keys.push_back((uint8_t)(h >> 8));
        }
    }
    st.keyCount += (uint32_t)K;
}

void pad4(std::vector<uint8_t>& v)
{
    while (v.size() % 4) v.push_back(0);
}

void emit_stream(std::vector<uint8_t>& out, Stream& st)
{
    out.insert(out.end(), st.desc.begin(), st.desc.end());
    pad4(st.bins);
    out.insert(out.end(), st.bins.begin(), st.bins.end());
    pad4(st.keys);
    out.insert(out.end(), st.keys.begin(), st.keys.end());
}

void put32(std::vector<uint8_t>& v, uint32_t x)
{
    for (int i = 0; i < 4; ++i) v.push_back((uint8_t)(x >> (8 * i)));
}
void put16(std::vector<uint8_t>& v, uint32_t x)
{
    v.push_back((uint8_t)(x & 0xFF));
    v.push_back((uint8_t)((x >> 8) & 0xFF));
}

float trilinear_lattice(const float* g, int res, double x, double y, double z)
{
    // lattice coordinate i * 7 / (res - 1) along each axis, positions in [0, 7]
    const double sc = (res - 1) / 7.0;
    double f[3] = {x * sc, y * sc, z * sc};
    int i0[3];
    double w[3];
    for (int a = 0; a < 3; ++a) {
        f[a] = std::min(std::max(f[a], 0.0), (double)(res - 1));
        i0[a] = std::min((int)std::floor(f[a]), res - 2);
        w[a] = f[a] - i0[a];
    }
    double v = 0.0;
    for (int c = 0; c < 8; ++c) {
        const int ix = i0[0] + (c & 1), iy = i0[1] + ((c >> 1) & 1), iz = i0[2] + ((c >> 2) & 1);
        const double wt = ((c & 1) ? w[0] : 1 - w[0]) * (((c >> 1) & 1) ? w[1] : 1 - w[1]) * (((c >> 2) & 1) ? w[2] : 1 - w[2]);
        v += wt * g[(iz * res + iy) * res + ix];
    }
    return (float)v;
}

struct LeafField {
    const Field& f;
    const std::vector<int>* sub;
    const vbts_params& params() const { return f.p; }
    double eval(double x, double y, double z, double t, bool voxel) const { return f.eval(x, y, z, t, voxel, sub); }
    bool leaf_may_be_nonzero(int x0, int y0, int z0) const
    {
        if (f.p.kind != VBTS_DENSITY && f.p.kind != VBTS_FLAMES) return true;
        return !sub->empty() && f.leaf_may_be_nonzero(x0, y0, z0);
    }
};

template <class FieldLike>
void controls_for(const FieldLike& fd, int res, int x0, int y0, int z0, int F, std::vector<float>& series)
{
    const int C = res * res * res;
    series.assign((size_t)C * F, 0.0f);
    const double sc = 7.0 / (res - 1);
    for (int c = 0; c < C; ++c) {
        const int ix = c % res, iy = (c / res) % res, iz = c / (res * res);
        for (int t = 0; t < F; ++t)
            series[(size_t)c * F + t] =
                (float)fd.eval(x0 + ix * sc, y0 + iy * sc, z0 + iz * sc, (double)t, false);
    }
}


// The series a control's keyframes decode to: the first key before it, the
// last after it, v0 + a (v1 - v0) between (fp16 keys rounded as stored).
void reconstruct(const float* s, const std::vector<int>& keys, int F, int codec, float* rec)
{
    auto val = [&](int k) { return codec == 1 ? s[k] : from_half(to_half(s[k])); };
    size_t i = 0;
    for (int t = 0; t < F; ++t) {
        while (i + 1 < keys.size() && keys[i + 1] <= t) ++i;
        if (t <= keys[i] || i + 1 >= keys.size()) {
            rec[t] = val(keys[i]);
            continue;
        }
        const int f0 = keys[i], f1 = keys[i + 1];
        const float a = (float)(t - f0) / (float)(f1 - f0);
        const float v0 = val(f0), v1 = val(f1);
        rec[t] = v0 + a * (v1 - v0);
    }
}

void build_stream(const std::vector<float>& series, int C, int F, int B, int codec, float epsScale,
                  const vbts_encode_opts& o, uint64_t seed, Stream& st, std::vector<float>* rec = nullptr)
{
    if (rec) rec->assign((size_t)C * F, 0.0f);
    std::vector<int> keys;
    for (int c = 0; c < C; ++c) {
        const float* s = &series[(size_t)c * F];
        fit_keys(s, F, KeyProfile{o.epsAbs * epsScale, o.epsRel * epsScale, o.gammaDelta, o.cutoff, o.cutoffBand},
                 keys);
        if (o.burstFraction > 0.0f && o.burstMax > 1) {
            const uint64_t h = mix64(seed ^ ((uint64_t)c * 0x9E3779B97F4A7C15ull));
            if (hash01(h) < (double)o.burstFraction) {
                const int extra = 1 + (int)(mix64(h) % (uint64_t)(o.burstMax - 1));
                std::vector<char> on(F, 0);
                for (int k : keys) on[k] = 1;
                uint64_t r = h;
                for (int e = 0; e < extra; ++e) {
                    r = mix64(r + 0x1234567ull);
                    on[(int)(r % (uint64_t)F)] = 1;
                }
                keys.clear();
                for (int t = 0; t < F; ++t)
                    if (on[t]) keys.push_back(t);
            }
        }
        append_control(st, s, keys, F, B, codec);
        if (rec) reconstruct(s, keys, F, codec, rec->data() + (size_t)c * F);
    }
}

// The range of a render field's values over its voxels and frames, which the
// route's errors are relative to (the reference divides by the raw data's
// max - min).  Density: the largest value at the voxels around every blob
// centre at every frame (the field is a sum of blobs and 0 far from them);
// flames peak at 1 (the band of the density); temperature spans 300 K to
// 1800 K.
double field_range(const Field& f)
{
    const vbts_params& p = f.p;
    if (p.kind == VBTS_FLAMES) return 1.0;
    if (p.kind == VBTS_TEMPERATURE) return 1500.0;
    if (p.kind != VBTS_DENSITY) return 1.0;
    double mx = 0.0;
    for (uint32_t t = 0; t < p.frames; ++t)
        for (size_t i = 0; i < f.blobs.size(); ++i) {
            double tmp[3];
            const double* c = f.centre(i, (double)t, tmp);
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const double x = std::floor(c[0] + 0.5) + dx, y = std::floor(c[1] + 0.5) + dy,
                                     z = std::floor(c[2] + 0.5) + dz;
                        if (x < 0 || y < 0 || z < 0 || x >= p.width || y >= p.height || z >= p.depth) continue;
                        mx = std::max(mx, f.eval(x, y, z, (double)t, true));
                    }
        }
    return std::max(1e-6, mx);
}

// leaf bytes: header, prefix (u16 keyframe counts), coarse stream, then the
// Fine6 stream or the Mode2 shell section (mask, group prefix, shell stream)
int emit_render_leaf(std::vector<uint8_t>& out, const vbts_encode_opts& o, int mode, Stream& cs, Stream& fs,
                     const std::vector<uint16_t>& shellVox)
{
    const uint32_t header = (uint32_t)mode | ((uint32_t)o.codec << 2) | ((uint32_t)o.codec << 4) | (4u << 6) |
                            ((mode == 3 ? 6u : 0u) << 9);
    put32(out, header);
    put16(out, cs.keyCount);
    put16(out, mode == 1 ? 0u : fs.keyCount);
    emit_stream(out, cs);
    if (mode == 3) emit_stream(out, fs);
    if (mode == 2) {
        uint64_t mask[8] = {0};
        for (uint16_t v : shellVox) mask[v >> 6] |= 1ull << (v & 63);
        for (int gI = 0; gI < 8; ++gI)
            for (int b = 0; b < 8; ++b) out.push_back((uint8_t)(mask[gI] >> (8 * b)));
        uint32_t prefix = 0;
        for (int gI = 0; gI < 8; ++gI) {
            put16(out, prefix);
            prefix += (uint32_t)__builtin_popcountll(mask[gI]);
        }
        emit_stream(out, fs);
    }
    return 0;
}

// Fine6 or Grid4 for a non-empty leaf of a blob field, by the reference's
// route rule (routeRenderTemporalLeaf, src/render_temporal_route.cpp:27-66, on
// metrics taken as render_temporal_kf_probe_main.cpp:1185-1240 takes them):
// at every voxel and frame of the leaf whose coordinates are multiples of 4,
// the field against the decoded coarse grid and against coarse + decoded 6^3
// residual, errors relative to the field's range.
bool route_fine6(const LeafField& fd, const vbts_encode_opts& o, double range, int x0, int y0, int z0, int lw,
                 int lh, int ld, int F, const std::vector<float>& recCoarse, const std::vector<float>& recFine)
{
    double ce2 = 0.0, fe2 = 0.0, cpk = 0.0;
    uint64_t n = 0, visible = 0, inBand = 0;
    const double lo = (double)o.cutoff - o.cutoffBand, hi = (double)o.cutoff + o.cutoffBand;
    std::vector<float> g(64), g6(216);
    for (int t = 0; t < F; t += 4) {
        for (int c = 0; c < 64; ++c) g[c] = recCoarse[(size_t)c * F + t];
        for (int c = 0; c < 216; ++c) g6[c] = recFine[(size_t)c * F + t];
        for (int z = (4 - z0 % 4) % 4; z < ld; z += 4)
            for (int y = (4 - y0 % 4) % 4; y < lh; y += 4)
                for (int x = (4 - x0 % 4) % 4; x < lw; x += 4) {
                    const double truth = fd.eval(x0 + x, y0 + y, z0 + z, t, true);
                    const double cp = trilinear_lattice(g.data(), 4, x, y, z);
                    const double fp = cp + trilinear_lattice(g6.data(), 6, x, y, z);
                    ++n;
                    visible += truth >= lo;
                    inBand += truth >= lo && truth <= hi;
                    ce2 += (truth - cp) * (truth - cp);
                    fe2 += (truth - fp) * (truth - fp);
                    cpk = std::max(cpk, std::fabs(truth - cp) / range);
                }
    }
    if (n == 0) return false;
    const double cr = std::sqrt(ce2 / n) / range, fr = std::sqrt(fe2 / n) / range;
    const bool rough = cr > o.routeCoarseRmse || cpk > o.routeCoarsePeak;
    const bool seen = (double)visible / n >= o.routeVisible || (double)inBand / n >= o.routeBand;
    const double gain = (cr - fr) / std::max(cr, 1e-12);
    return rough && seen && gain >= o.routeFineGain;
}

// A stream's keyframe count is u16 in the leaf prefix: its tolerance grows by
// 1.25x until it fits.
int build_stream_fit(const std::vector<float>& series, int C, int F, int B, const vbts_encode_opts& o, uint64_t seed,
                     Stream& st, std::vector<float>* rec)
{
    for (float sc = 1.0f;; sc *= 1.25f) {
        st = Stream{};
        build_stream(series, C, F, B, o.codec, sc, o, seed, st, rec);
        if (st.keyCount <= 65535u) return 0;
        if (sc > 1e6f) {
            g_err = "leaf keyframe count exceeds u16";
            return 1;
        }
    }
}

// The Mode2 shell stream: each shell voxel's residual keyed with the iso keys
// of its series; the tolerance grows by 1.25x until the keyframes fit u16.
int build_shell_stream(const std::vector<float>& resid, const std::vector<float>& truth, int C, int F, float isoBand,
                       const vbts_encode_opts& o, uint64_t seed, Stream& st)
{
    (void)seed;
    std::vector<int> keys;
    for (float sc = 1.0f;; sc *= 1.25f) {
        st = Stream{};
        const KeyProfile kp{o.epsAbs * sc, o.epsRel * sc, o.gammaDelta, o.cutoff, o.cutoffBand};
        for (int c = 0; c < C; ++c) {
            fit_keys(&resid[(size_t)c * F], F, kp, keys, &truth[(size_t)c * F], isoBand);
            append_control(st, &resid[(size_t)c * F], keys, F, 4, o.codec);
        }
        if (st.keyCount <= 65535u) return 0;
        if (sc > 1e6f) {
            g_err = "leaf keyframe count exceeds u16";
            return 1;
        }
    }
}

int encode_render_leaf(const Field& full, const vbts_encode_opts& o, double range, uint32_t leaf,
                       std::vector<uint8_t>& out)
{
    const vbts_params& p = full.p;
    const int lcx = (int)((p.width + 7) / 8), lcy = (int)((p.height + 7) / 8);
    const int bx = (int)(leaf % lcx), by = (int)((leaf / lcx) % lcy), bz = (int)(leaf / (lcx * lcy));
    const int x0 = bx * 8, y0 = by * 8, z0 = bz * 8;
    // the field restricted to the items that can reach this leaf
    const std::vector<int> sub = leaf_subset(full, x0, y0, z0);
    const LeafField fd{full, &sub};
    const int lw = std::min(8, (int)p.width - x0), lh = std::min(8, (int)p.height - y0), ld = std::min(8, (int)p.depth - z0);
    const int F = (int)p.frames;
    out.clear();
    if (!fd.leaf_may_be_nonzero(x0, y0, z0)) {
        put32(out, 0u);
        return 0;
    }
    std::vector<float> coarse;
    controls_for(fd, 4, x0, y0, z0, F, coarse);
    bool allZero = true;
    for (float v : coarse) if (v != 0.0f) { allZero = false; break; }
    if (allZero && (p.kind == VBTS_DENSITY || p.kind == VBTS_FLAMES)) {
        // probe voxels (every 2nd voxel, every 4th frame) before calling the leaf empty
        bool any = false;
        for (int t = 0; t < F && !any; t += 4)
            for (int z = 0; z < ld && !any; z += 2)
                for (int y = 0; y < lh && !any; y += 2)
                    for (int x = 0; x < lw && !any; x += 2)
                        if (fd.eval(x0 + x, y0 + y, z0 + z, t, true) != 0.0) any = true;
        if (!any) {
            put32(out, 0u);
            return 0;
        }
    }
    const uint64_t seed = mix64(p.seed ^ ((uint64_t)leaf * 0x2545F4914F6CDD1Dull));
    // The fine / shell residuals are taken against the coarse grid as it
    // decodes (closed loop), so the coarse keyframe error does not add to theirs.
    Stream cs, fs;
    std::vector<float> recCoarse;
    if (build_stream_fit(coarse, 64, F, 8, o, seed, cs, &recCoarse)) return 1;
    int mode = 1; // GRID4
    if (p.kind == VBTS_LEVELSET) {
        // Mode2 as the reference's level-set surface route builds it
        // (render_temporal_kf_probe_main.cpp:600-660, 940-1010): each voxel's
        // series keyed with iso-surface protection and replaced by what its
        // keys interpolate; a voxel is a shell voxel where that series comes
        // within the surface band of 0, the decoded coarse grid within the
        // guard band, or the two differ in sign, at some frame; its series is
        // the residual to the coarse grid, keyed with its iso keys
        const float surface = o.shellBand, guard = std::max(0.5f, 0.5f * o.shellBand), isoBand = 0.1f;
        // a voxel whose values all stay below this is background (its series
        // becomes 0), a leaf of such voxels is empty (bgZeroRatio 0.3 of the cutoff)
        const float bg = 0.3f * o.cutoff;
        const KeyProfile vp{o.epsAbs, o.epsRel, o.gammaDelta, o.cutoff, o.cutoffBand};
        std::vector<uint16_t> shellVox;
        std::vector<float> truth, resid;
        double minAbs = 1e30, maxCentre = -1e30;
        for (int t = 0; t < F; ++t) {
            const double c = fd.eval(x0 + 3.5, y0 + 3.5, z0 + 3.5, t, false);
            minAbs = std::min(minAbs, std::fabs(c));
            maxCentre = std::max(maxCentre, c);
        }
        // phi changes by about a voxel per voxel: a centre this far inside at
        // every frame leaves every voxel of the leaf below the background
        if (maxCentre < -((double)surface + 6.1)) {
            out.clear();
            put32(out, 0u);
            return 0;
        }
        if (minAbs < (double)surface + 6.1) {
            float leafMax = -1e30f;
            std::vector<float> g(64), raw(F), rec(F), tr((size_t)512 * F), cg((size_t)512 * F);
            std::vector<int> keys;
            for (int t = 0; t < F; ++t) {
                for (int c = 0; c < 64; ++c) g[c] = recCoarse[(size_t)c * F + t];
                for (int v = 0; v < 512; ++v)
                    cg[(size_t)v * F + t] = trilinear_lattice(g.data(), 4, v & 7, (v >> 3) & 7, v >> 6);
            }
            for (int z = 0; z < ld; ++z)
                for (int y = 0; y < lh; ++y)
                    for (int x = 0; x < lw; ++x) {
                        const int v = (z * 8 + y) * 8 + x;
                        float vmax = -1e30f;
                        for (int t = 0; t < F; ++t) {
                            raw[t] = (float)fd.eval(x0 + x, y0 + y, z0 + z, t, true);
                            vmax = std::max(vmax, raw[t]);
                        }
                        if (vmax < bg) {
                            std::fill(rec.begin(), rec.end(), 0.0f);
                        } else {
                            fit_keys(raw.data(), F, vp, keys, raw.data(), isoBand);
                            recon_linear(raw.data(), keys, F, rec.data());
                        }
                        for (int t = 0; t < F; ++t) leafMax = std::max(leafMax, rec[t]);
                        const float* c = &cg[(size_t)v * F];
                        bool occ = false;
                        for (int t = 0; t < F && !occ; ++t)
                            occ = std::fabs(rec[t]) <= surface || std::fabs(c[t]) <= guard || (rec[t] >= 0.0f) != (c[t] >= 0.0f);
                        if (!occ) continue;
                        shellVox.push_back((uint16_t)v);
                        std::copy(rec.begin(), rec.end(), tr.begin() + (size_t)v * F);
                    }
            if (leafMax < bg) {
                out.clear();
                put32(out, 0u);
                return 0;
            }
            if (!shellVox.empty()) {
                mode = 2;
                resid.resize(shellVox.size() * (size_t)F);
                truth.resize(shellVox.size() * (size_t)F);
                for (size_t i = 0; i < shellVox.size(); ++i)
                    for (int t = 0; t < F; ++t) {
                        const size_t vt = (size_t)shellVox[i] * F + t;
                        truth[i * F + t] = tr[vt];
                        resid[i * F + t] = tr[vt] - cg[vt];
                    }
            }
        }
        if (mode == 2 && build_shell_stream(resid, truth, (int)shellVox.size(), F, isoBand, o, seed ^ 2u, fs)) return 1;
        return emit_render_leaf(out, o, mode, cs, fs, shellVox);
    }
    // the 6^3 residual of the decoded coarse grid, kept if the route picks Fine6
    std::vector<float> resid((size_t)216 * F), recFine;
    {
        std::vector<float> g(64);
        const double sc6 = 7.0 / 5.0;
        for (int t = 0; t < F; ++t) {
            for (int c = 0; c < 64; ++c) g[c] = recCoarse[(size_t)c * F + t];
            for (int j = 0; j < 216; ++j) {
                const double x = (j % 6) * sc6, y = ((j / 6) % 6) * sc6, z = (j / 36) * sc6;
                resid[(size_t)j * F + t] =
                    (float)fd.eval(x0 + x, y0 + y, z0 + z, t, false) - trilinear_lattice(g.data(), 4, x, y, z);
            }
        }
    }
    if (build_stream_fit(resid, 216, F, 8, o, seed ^ 1u, fs, &recFine)) return 1;
    if (route_fine6(fd, o, range, x0, y0, z0, lw, lh, ld, F, recCoarse, recFine)) mode = 3;
    return emit_render_leaf(out, o, mode, cs, fs, {});
}

// ---------------------------------------------------------- scientific --
double dct_basis_d(int N, int t, int k)
{
    const double invN = 1.0 / (double)N;
    const double alpha = k == 0 ? std::sqrt(invN) : std::sqrt(2.0 * invN);
    return alpha * std::cos((3.14159265358979323846 / (double)N) * ((double)t + 0.5) * (double)k);
}

float dct_basis_f(int N, int t, int k) { return (float)dct_basis_d(N, t, k); }

int encode_scientific(const Field& fd, const vbts_encode_opts& o, std::vector<uint8_t>& file)
{
    const vbts_params& p = fd.p;
    const int F = (int)p.frames;
    const int keep = std::max(1, std::min({o.sciKeep, 64, F}));
    const uint32_t lcx = (p.width + 7) / 8, lcy = (p.height + 7) / 8, lcz = (p.depth + 7) / 8;
    const uint32_t L = lcx * lcy * lcz;
    std::vector<double> coef((size_t)L * 64 * keep);
    const int threads = o.threads > 0 ? o.threads : 1;
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads)
    for (int64_t l = 0; l < (int64_t)L; ++l) {
        const int x0 = (int)(l % lcx) * 8, y0 = (int)((l / lcx) % lcy) * 8, z0 = (int)(l / (lcx * lcy)) * 8;
        std::vector<double> s(F);
        for (int c = 0; c < 64; ++c) {
            const int cx = c & 3, cy = (c >> 2) & 3, cz = c >> 4;
            for (int t = 0; t < F; ++t)
                s[t] = fd.smooth_sci(x0 + cx * 7.0 / 3.0, y0 + cy * 7.0 / 3.0, z0 + cz * 7.0 / 3.0, t);
            for (int k = 0; k < keep; ++k) {
                double a = 0.0;
                for (int t = 0; t < F; ++t) a += s[t] * dct_basis_d(F, t, k);
                coef[((size_t)l * 64 + c) * keep + k] = a;
            }
        }
    }
    std::vector<float> acScale(keep > 1 ? keep - 1 : 0, 1e-12f);
    for (size_t i = 0; i < coef.size(); ++i) {
        const int k = (int)(i % keep);
        if (k > 0) acScale[k - 1] = std::max(acScale[k - 1], (float)(std::fabs(coef[i]) / 127.0));
    }
    std::vector<float> basisF((size_t)F * keep);
    for (int t = 0; t < F; ++t)
        for (int k = 0; k < keep; ++k) basisF[(size_t)t * keep + k] = dct_basis_f(F, t, k);
    const int fpb = (F + 15) / 16;
    if (fpb > 16) {
        g_err = "scientific sparse events need frames <= 256";
        return 1;
    }
    std::vector<std::vector<uint8_t>> leaves(L);
    std::atomic<int> err{0};
#pragma omp parallel for schedule(dynamic, 2) num_threads(threads)
    for (int64_t l = 0; l < (int64_t)L; ++l) {
        const int x0 = (int)(l % lcx) * 8, y0 = (int)((l / lcx) % lcy) * 8, z0 = (int)(l / (lcx * lcy)) * 8;
        const int lw = std::min(8, (int)p.width - x0), lh = std::min(8, (int)p.height - y0), ld = std::min(8, (int)p.depth - z0);
        std::vector<uint8_t> body;
        std::vector<uint16_t> dc(64);
        std::vector<int8_t> ac((size_t)64 * keep);
        for (int c = 0; c < 64; ++c) {
            const double* a = &coef[((size_t)l * 64 + c) * keep];
            dc[c] = to_half((float)a[0]);
            body.push_back((uint8_t)(dc[c] & 0xFF));
            body.push_back((uint8_t)(dc[c] >> 8));
            for (int k = 1; k < keep; ++k) {
                const double q = std::round(a[k] / acScale[k - 1]);
                ac[(size_t)c * keep + k] = (int8_t)std::max(-127.0, std::min(127.0, q));
                body.push_back((uint8_t)ac[(size_t)c * keep + k]);
            }
        }
        // decoded coarse controls per frame (decoder arithmetic), then residual at voxels
        std::vector<float> resid((size_t)F * 512, 0.0f);
        double rms = 0.0;
        float ctrl[64];
        for (int t = 0; t < F; ++t) {
            for (int c = 0; c < 64; ++c) {
                double sum = (double)from_half(dc[c]) * (double)basisF[(size_t)t * keep];
                for (int k = 1; k < keep; ++k) {
                    const float coeff = (float)ac[(size_t)c * keep + k] * acScale[k - 1];
                    sum += (double)coeff * (double)basisF[(size_t)t * keep + k];
                }
                ctrl[c] = (float)sum;
            }
            for (int z = 0; z < ld; ++z)
                for (int y = 0; y < lh; ++y)
                    for (int x = 0; x < lw; ++x) {
                        const double v = fd.eval(x0 + x, y0 + y, z0 + z, t, true);
                        const float cv = trilinear_lattice(ctrl, 4, x, y, z);
                        const float r = (float)(v - cv);
                        resid[(size_t)t * 512 + (z * 8 + y) * 8 + x] = r;
                        rms += (double)r * r;
                    }
        }
        rms = std::sqrt(rms / std::max(1, F * lw * lh * ld));
        // sparse events: residual outliers (spikes)
        struct Ev { uint32_t bin; uint16_t key; float r; };
        std::vector<Ev> ev;
        const float spikeThr = (float)(0.5 * p.spikeGain * fd.amp);
        for (int t = 0; t < F; ++t)
            for (int v = 0; v < 512; ++v) {
                const float r = resid[(size_t)t * 512 + v];
                if (p.spikeGain > 0 && std::fabs(r) > spikeThr) {
                    const uint32_t bin = std::min(15u, (uint32_t)(t / fpb));
                    const uint32_t lt = (uint32_t)t - bin * (uint32_t)fpb;
                    ev.push_back({bin, (uint16_t)((lt << 9) | (uint32_t)v), r});
                }
            }
        uint32_t header = 0;
        std::vector<uint8_t> tail;
        const uint64_t lh64 = mix64(p.seed ^ (uint64_t)l);
        if (!ev.empty()) {
            if (ev.size() > 1024) {
                std::nth_element(ev.begin(), ev.begin() + 1024, ev.end(),
                                 [](const Ev& a, const Ev& b) { return std::fabs(a.r) > std::fabs(b.r); });
                ev.resize(1024);
            }
            std::sort(ev.begin(), ev.end(), [](const Ev& a, const Ev& b) {
                return a.bin != b.bin ? a.bin < b.bin : a.key < b.key;
            });
            const uint32_t n = (uint32_t)ev.size();
            const uint32_t tier = n <= 64 ? 0u : n <= 256 ? 1u : n <= 768 ? 2u : 3u;
            const uint32_t cap = tier == 0 ? 64u : tier == 1 ? 256u : tier == 2 ? 768u : 1024u;
            float mx = 0.0f;
            for (const auto& e : ev) mx = std::max(mx, std::fabs(e.r));
            const float es = mx / 7.5f;
            uint32_t esb;
            std::memcpy(&esb, &es, 4);
            put32(tail, esb);
            uint32_t mask[16] = {0};
            for (const auto& e : ev) mask[(e.key & 0x1FFu) >> 5] |= 1u << (e.key & 31u);
            for (int i = 0; i < 16; ++i) put32(tail, mask[i]);
            for (uint32_t b = 0, i = 0; b < 16; ++b) {
                while (i < n && ev[i].bin < b) ++i;
                put16(tail, i);
            }
            for (uint32_t i = 0; i < cap; ++i) put16(tail, i < n ? ev[i].key : 0xFFFFu);
            std::vector<uint8_t> nib((cap + 1) / 2, 0);
            for (uint32_t i = 0; i < n; ++i) {
                int q = (int)std::lround(ev[i].r / es);
                q = std::max(-8, std::min(7, q));
                nib[i >> 1] |= (uint8_t)(((uint32_t)q & 0xFu) << ((i & 1) * 4));
            }
            tail.insert(tail.end(), nib.begin(), nib.end());
            header = 1u | ((uint32_t)(keep - 1) << 2) | (tier << 8) | ((uint32_t)(fpb - 1) << 10) | (n << 16);
        } else if (rms > 0.004 * fd.amp) {
            const int variant = (int)(lh64 % 4u);
            const int R = (variant & 1) ? 3 : 4;
            const int sub = (variant >> 1) & 1;
            const int bits = R == 3 ? 3 : 4;
            const int levels = (1 << bits) - 1;
            const int tk = 4;
            const int vals = R * R * R;
            const int blockBytes = (vals * bits + 7) / 8;
            // residual on the R-lattice: nearest voxel of lattice coordinate c * 7 / (R - 1)
            auto grid_at = [&](int t, int c) {
                const int cx = c % R, cy = (c / R) % R, cz = c / (R * R);
                int x = (int)std::lround(cx * 7.0 / (R - 1)), y = (int)std::lround(cy * 7.0 / (R - 1)),
                    z = (int)std::lround(cz * 7.0 / (R - 1));
                x = std::min(x, lw - 1); y = std::min(y, lh - 1); z = std::min(z, ld - 1);
                return resid[(size_t)t * 512 + (z * 8 + y) * 8 + x];
            };
            const int nb = sub == 0 ? F : tk;
            std::vector<float> blocks((size_t)nb * vals);
            for (int b = 0; b < nb; ++b)
                for (int c = 0; c < vals; ++c) {
                    if (sub == 0) blocks[(size_t)b * vals + c] = grid_at(b, c);
                    else {
                        double a = 0.0;
                        for (int t = 0; t < F; ++t) a += grid_at(t, c) * dct_basis_d(F, t, b);
                        blocks[(size_t)b * vals + c] = (float)a;
                    }
                }
            std::vector<uint8_t> codes;
            for (int b = 0; b < nb; ++b) {
                float mx = 0.0f;
                for (int c = 0; c < vals; ++c) mx = std::max(mx, std::fabs(blocks[(size_t)b * vals + c]));
                const uint16_t hs = to_half(mx);
                put16(tail, hs);
                const float sc = from_half(hs);
                std::vector<uint8_t> blk(blockBytes, 0);
                for (int c = 0; c < vals; ++c) {
                    const float v = sc > 0.0f ? blocks[(size_t)b * vals + c] / sc : 0.0f;
                    int q = (int)std::lround((v + 1.0f) * 0.5f * levels);
                    q = std::max(0, std::min(levels, q));
                    const int bit = c * bits;
                    const uint32_t w = (uint32_t)q << (bit % 8);
                    blk[bit / 8] |= (uint8_t)(w & 0xFF);
                    if (bit % 8 + bits > 8) blk[bit / 8 + 1] |= (uint8_t)(w >> 8);
                }
                codes.insert(codes.end(), blk.begin(), blk.end());
            }
            tail.insert(tail.end(), codes.begin(), codes.end());
            header = (R == 3 ? 2u : 3u) | ((uint32_t)(keep - 1) << 2) | ((uint32_t)bits << 8) | ((uint32_t)sub << 12) |
                     ((uint32_t)(tk - 1) << 15);
        } else {
            header = 0u | ((uint32_t)(keep - 1) << 2);
        }
        std::vector<uint8_t>& out = leaves[l];
        put32(out, header);
        out.insert(out.end(), body.begin(), body.end());
        out.insert(out.end(), tail.begin(), tail.end());
        pad4(out);
    }
    if (err.load()) return 1;
    // container
    file.clear();
    const char magic[8] = {'V', 'B', 'T', 'P', 'A', 'C', 'K', '4'};
    file.insert(file.end(), magic, magic + 8);
    const uint32_t hdr[11] = {4u, p.width, p.height, p.depth, p.frames, 8u, 4u, (uint32_t)keep, L, 2u, (uint32_t)(keep - 1)};
    for (uint32_t v : hdr) put32(file, v);
    for (float s : acScale) {
        uint32_t b;
        std::memcpy(&b, &s, 4);
        put32(file, b);
    }
    uint64_t words = 0;
    put32(file, 0u);
    for (uint32_t l = 0; l < L; ++l) {
        words += leaves[l].size() / 4;
        if (words > 0xffffffffull) {
            g_err = "payload exceeds u32 word offsets";
            return 1;
        }
        put32(file, (uint32_t)words);
    }
    for (uint32_t l = 0; l < L; ++l) file.insert(file.end(), leaves[l].begin(), leaves[l].end());
    return 0;
}

} // namespace

extern "C" {

const char* vbts_last_error(void) { return g_err.c_str(); }

void vbts_default_params(int kind, uint32_t w, uint32_t h, uint32_t d, uint32_t f, uint64_t seed, vbts_params* p)
{
    std::memset(p, 0, sizeof(*p));
    p->kind = kind;
    p->width = w;
    p->height = h;
    p->depth = d;
    p->frames = f;
    p->seed = seed;
    p->count = 6;
    p->sizeMin = 4.0f;
    p->sizeMax = 10.0f;
    p->peakMin = 0.2f;
    p->peakMax = 1.0f;
    p->noise = 0.01f;
    p->motion = 1.0f;
    p->spikeFraction = 0.001f;
    p->spikeGain = 50.0f;
    p->support = 2.0f;
    if (kind == VBTS_LEVELSET) {
        p->count = 40;
        p->sizeMin = 4.0f;
        p->sizeMax = 24.0f;
        p->noise = 0.0f;
    } else if (kind == VBTS_SCIENTIFIC) {
        p->count = 6;
        p->noise = 0.0f;
    }
}

void* vbts_field_create(const vbts_params* p)
{
    auto* f = new Field();
    f->p = *p;
    build_field(*f);
    return f;
}

void vbts_field_free(void* field) { delete static_cast<Field*>(field); }

float vbts_field_eval(void* field, float x, float y, float z, uint32_t t)
{
    return (float)static_cast<Field*>(field)->eval(x, y, z, t, true);
}

int vbts_field_frame(void* field, uint32_t t, float* out, int threads)
{
    const Field& f = *static_cast<Field*>(field);
    const int64_t W = f.p.width, H = f.p.height, D = f.p.depth;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
    for (int64_t z = 0; z < D; ++z)
        for (int64_t y = 0; y < H; ++y)
            for (int64_t x = 0; x < W; ++x)
                out[(z * H + y) * W + x] = (float)f.eval((double)x, (double)y, (double)z, (double)t, true);
    return 0;
}

void vbts_default_encode_opts(vbts_encode_opts* o)
{
    std::memset(o, 0, sizeof(*o));
    o->codec = 0;
    // the reference render encoder's defaults (ProbeCliOptions,
    // render_temporal_kf_probe_main.cpp:37-71)
    o->epsAbs = 1e-5f;
    o->epsRel = 0.02f;
    o->gammaDelta = 0.2f;
    o->cutoff = 0.0003f;
    o->cutoffBand = 0.0001f;
    o->routeCoarseRmse = 0.015f;
    o->routeCoarsePeak = 0.06f;
    o->routeVisible = 0.02f;
    o->routeBand = 0.01f;
    o->routeFineGain = 0.15f;
    o->shellBand = 1.5f; // the reference's level-set surface band without shell_width_voxels
    o->burstFraction = 0.0f;
    o->burstMax = 32;
    o->sciKeep = 12;
    o->threads = 8;
}

int vbts_encode(void* field, const vbts_encode_opts* o, uint8_t** outBytes, uint64_t* outSize)
{
    const Field& fd = *static_cast<Field*>(field);
    const vbts_params& p = fd.p;
    std::vector<uint8_t> file;
    if (p.frames == 0 || p.width == 0 || p.height == 0 || p.depth == 0) {
        g_err = "empty field";
        return 1;
    }
    if (p.kind == VBTS_SCIENTIFIC) {
        if (encode_scientific(fd, *o, file)) return 1;
    } else {
        if (o->codec == 0 && p.frames > 256) {
            g_err = "fp16 keyframes store frames as u8 (frames <= 256)";
            return 1;
        }
        const uint32_t lcx = (p.width + 7) / 8, lcy = (p.height + 7) / 8, lcz = (p.depth + 7) / 8;
        const uint32_t L = lcx * lcy * lcz;
        const double range = field_range(fd);
        std::vector<std::vector<uint8_t>> leaves(L);
        std::atomic<int> err{0};
        std::string msg;
        std::mutex mu;
#pragma omp parallel for schedule(dynamic, 4) num_threads(o->threads > 0 ? o->threads : 1)
        for (int64_t l = 0; l < (int64_t)L; ++l) {
            if (err.load()) continue;
            if (encode_render_leaf(fd, *o, range, (uint32_t)l, leaves[l])) {
                std::lock_guard<std::mutex> g(mu);
                err = 1;
                msg = g_err;
            }
        }
        if (err.load()) {
            g_err = msg;
            return 1;
        }
        const char magic[8] = {'V', 'B', 'T', 'P', 'A', 'C', 'K', '4'};
        file.insert(file.end(), magic, magic + 8);
        const uint32_t hdr[11] = {4u, p.width, p.height, p.depth, p.frames, 8u, 4u, 0u, L, 1u, 0u};
        for (uint32_t v : hdr) put32(file, v);
        uint64_t words = 0;
        put32(file, 0u);
        for (uint32_t l = 0; l < L; ++l) {
            words += leaves[l].size() / 4;
            if (words > 0xffffffffull) {
                g_err = "payload exceeds u32 word offsets";
                return 1;
            }
            put32(file, (uint32_t)words);
        }
        file.reserve(file.size() + words * 4);
        for (uint32_t l = 0; l < L; ++l) file.insert(file.end(), leaves[l].begin(), leaves[l].end());
    }
    *outBytes = (uint8_t*)std::malloc(file.size());
    if (!*outBytes) {
        g_err = "out of memory";
        return 1;
    }
    std::memcpy(*outBytes, file.data(), file.size());
    *outSize = file.size();
    return 0;
}

void vbts_free(void* p) { std::free(p); }

} // extern "C"
