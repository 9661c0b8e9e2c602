/*
 * Seeded synthetic "flat-shaded" content for benchmarks and tests: a uniform
 * background, 6-12 filled polygons of uniform level, 1-3 dark cubic Bezier strokes
 * of width 1-3 px, additive Gaussian noise (sigma 0.01), clamped to [0,1].
 *
 * Deterministic on any platform: the only randomness is xoshiro256** (Blackman &
 * Vigna, "Scrambled linear pseudorandom number generators", 2018) seeded through
 * splitmix64, normal deviates by the Box-Muller transform, and the geometry uses
 * only +,-,*,/, sqrt, sin and cos of doubles (libm's correctly rounded sqrt; sin/cos
 * only place polygon vertices, so a last-bit libm difference moves an edge by far
 * less than a pixel).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { uint64_t s[4]; } rng_t;

static uint64_t splitmix64(uint64_t* x)
{
    uint64_t z = (*x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static uint64_t next_u64(rng_t* r)
{
    uint64_t* s = r->s;
    uint64_t result = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

static void rng_seed(rng_t* r, uint64_t seed)
{
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
}

static double uniform(rng_t* r) { return (double)(next_u64(r) >> 11) * (1.0 / 9007199254740992.0); }
static int randint(rng_t* r, int lo, int hi) { return lo + (int)(next_u64(r) % (uint64_t)(hi - lo + 1)); }

static double gauss(rng_t* r, double* spare, int* has)
{
    if (*has) { *has = 0; return *spare; }
    double u1 = uniform(r), u2 = uniform(r);
    if (u1 < 1e-300) u1 = 1e-300;
    double mag = sqrt(-2.0 * log(u1));
    *spare = mag * sin(6.283185307179586 * u2);
    *has = 1;
    return mag * cos(6.283185307179586 * u2);
}

/* even-odd scanline fill of a closed polygon at pixel centres */
static void fill_polygon(float* img, int w, int h, int ch, const double* vx, const double* vy, int nv, const float* level)
{
    double ymin = vy[0], ymax = vy[0];
    for (int i = 1; i < nv; i++) { if (vy[i] < ymin) ymin = vy[i]; if (vy[i] > ymax) ymax = vy[i]; }
    int y0 = (int)floor(ymin - 0.5), y1 = (int)ceil(ymax - 0.5);
    if (y0 < 0) y0 = 0;
    if (y1 > h - 1) y1 = h - 1;
    double xs[16];
    for (int y = y0; y <= y1; y++) {
        double cy = y + 0.5;
        int n = 0;
        for (int i = 0; i < nv; i++) {
            int j = (i + 1) % nv;
            double ay = vy[i], by = vy[j];
            if ((ay <= cy && by > cy) || (by <= cy && ay > cy)) {
                double t = (cy - ay) / (by - ay);
                xs[n++] = vx[i] + t * (vx[j] - vx[i]);
            }
        }
        for (int a = 1; a < n; a++) {           /* insertion sort, n <= 16 */
            double v = xs[a];
            int b = a - 1;
            while (b >= 0 && xs[b] > v) { xs[b + 1] = xs[b]; b--; }
            xs[b + 1] = v;
        }
        for (int k = 0; k + 1 < n; k += 2) {
            int xa = (int)ceil(xs[k] - 0.5), xb = (int)ceil(xs[k + 1] - 0.5);   /* [xa, xb) */
            if (xa < 0) xa = 0;
            if (xb > w) xb = w;
            for (int x = xa; x < xb; x++)
                for (int c = 0; c < ch; c++) img[((size_t)y * w + x) * ch + c] = level[c];
        }
    }
}

static void stamp_disc(float* img, int w, int h, int ch, double x, double y, double rad, const float* level)
{
    int xa = (int)floor(x - rad), xb = (int)ceil(x + rad), ya = (int)floor(y - rad), yb = (int)ceil(y + rad);
    double r2 = rad * rad;
    int hit = 0;
    for (int py = ya; py <= yb; py++) {
        if (py < 0 || py >= h) continue;
        for (int px = xa; px <= xb; px++) {
            if (px < 0 || px >= w) continue;
            double dx = px + 0.5 - x, dy = py + 0.5 - y;
            if (dx * dx + dy * dy <= r2) {
                for (int c = 0; c < ch; c++) img[((size_t)py * w + px) * ch + c] = level[c];
                hit = 1;
            }
        }
    }
    if (!hit) {
        int px = (int)floor(x), py = (int)floor(y);
        if (px >= 0 && px < w && py >= 0 && py < h)
            for (int c = 0; c < ch; c++) img[((size_t)py * w + px) * ch + c] = level[c];
    }
}

/* out: w*h*ch floats in [0,1], interleaved channels */
int ac_synth_flat_shaded(uint64_t seed, int w, int h, int ch, float* out)
{
    if (w <= 0 || h <= 0 || ch <= 0 || ch > 4 || !out) return -22;
    rng_t r;
    rng_seed(&r, seed);
    float level[4];
    for (int c = 0; c < ch; c++) level[c] = (float)uniform(&r);
    for (size_t i = 0; i < (size_t)w * h; i++)
        for (int c = 0; c < ch; c++) out[i * ch + c] = level[c];

    const double side = w < h ? w : h;
    int npoly = randint(&r, 6, 12);
    for (int p = 0; p < npoly; p++) {
        double cx = uniform(&r) * w, cy = uniform(&r) * h;
        double rad = (0.05 + 0.45 * uniform(&r)) * side;
        int nv = randint(&r, 3, 8);
        double vx[8], vy[8];
        for (int i = 0; i < nv; i++) {
            double a = 6.283185307179586 * (i + 0.8 * uniform(&r)) / nv;
            double rr = rad * (0.35 + 0.65 * uniform(&r));
            vx[i] = cx + rr * cos(a);
            vy[i] = cy + rr * sin(a);
        }
        for (int c = 0; c < ch; c++) level[c] = (float)uniform(&r);
        fill_polygon(out, w, h, ch, vx, vy, nv, level);
    }
    int nstroke = randint(&r, 1, 3);
    for (int s = 0; s < nstroke; s++) {
        double px[4], py[4];
        for (int i = 0; i < 4; i++) { px[i] = uniform(&r) * w; py[i] = uniform(&r) * h; }
        double width = 1.0 + 2.0 * uniform(&r);
        for (int c = 0; c < ch; c++) level[c] = (float)(0.15 * uniform(&r));
        double len = 0.0;
        for (int i = 0; i < 3; i++) len += sqrt((px[i + 1] - px[i]) * (px[i + 1] - px[i]) + (py[i + 1] - py[i]) * (py[i + 1] - py[i]));
        int n = (int)(2.0 * len) + 2;
        for (int k = 0; k <= n; k++) {
            double t = (double)k / n, u = 1.0 - t;
            double b0 = u * u * u, b1 = 3 * u * u * t, b2 = 3 * u * t * t, b3 = t * t * t;
            double x = b0 * px[0] + b1 * px[1] + b2 * px[2] + b3 * px[3];
            double y = b0 * py[0] + b1 * py[1] + b2 * py[2] + b3 * py[3];
            stamp_disc(out, w, h, ch, x, y, 0.5 * width, level);
        }
    }
    double spare = 0.0;
    int has = 0;
    for (size_t i = 0; i < (size_t)w * h * ch; i++) {
        double v = out[i] + 0.01 * gauss(&r, &spare, &has);
        out[i] = (float)(v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v));
    }
    return 0;
}

/* u8 quantisation of the generator: round(255 v) */
int ac_synth_flat_shaded_u8(uint64_t seed, int w, int h, int ch, uint8_t* out)
{
    float* tmp = (float*)malloc(sizeof(float) * (size_t)w * h * (ch > 0 ? ch : 1));
    if (!tmp) return -12;
    int rc = ac_synth_flat_shaded(seed, w, h, ch, tmp);
    if (rc == 0)
        for (size_t i = 0; i < (size_t)w * h * ch; i++) out[i] = (uint8_t)(tmp[i] * 255.0f + 0.5f);
    free(tmp);
    return rc;
}

/* uniform integers in [lo, hi] from the same generator (image sizes of batched loads) */
int ac_synth_randint(uint64_t seed, int n, int lo, int hi, int* out)
{
    if (n < 0 || hi < lo || !out) return -22;
    rng_t r;
    rng_seed(&r, seed);
    for (int i = 0; i < n; i++) out[i] = randint(&r, lo, hi);
    return 0;
}

/* uniform doubles in [0,1) (arrival times of synthetic request streams) */
int ac_synth_uniform(uint64_t seed, int n, double* out)
{
    if (n < 0 || !out) return -22;
    rng_t r;
    rng_seed(&r, seed);
    for (int i = 0; i < n; i++) out[i] = uniform(&r);
    return 0;
}
