/* CPU oracle of the atmosphere lookup tables (test infrastructure only: the
 * tests, the package's smoke check and the bench's CPU leg load it; the
 * product never does).
 *
 * A restatement in FP32 of atmosphere.slang (Hillaire 2020, units in
 * megameters): transmittance 256x64 (transmittance_lut_compute,
 * atmosphere.slang:143-166), multiple scattering 32x32 (multiscatt_lut_compute
 * and getMulScattValues, :172-258) and sky-view 200x200
 * (sky_view_lut_compute and raymarchScattering, :262-397), plus the
 * transmittance seen from the view point (scene.slang:997-1009).  Images are
 * float4 rows [y][x].  LUT reads are bilinear with clamp-to-edge addressing.
 *
 * The GPU computes exp/sin/cos/acos/pow with its own math library, so LUT
 * values agree with this file to a relative tolerance, not bit for bit.
 *
 * Pinned by: scattering coefficients at closed-form altitudes; ray/sphere
 * distances in closed form; the vertical-path transmittance against the
 * geometric-series sum of the exponential densities; the horizon boundary of
 * the transmittance table; monotonicity in sun angle and height; the
 * azimuthal mirror symmetry of the sky-view table.  The multiple-scattering
 * table is pinned by invariants (finite and non-negative, exactly linear in
 * the transmittance table, zero where the sun is under the ground) and by
 * the integrals it discretises: the same code at 128 x 128 directions and 800
 * steps agrees with an independent float64 evaluation of Hillaire's L2 and
 * f_ms integrals within 1% (3% at grazing sun); the sky-view table is exactly
 * linear in its two input tables jointly and the sum of its single- and
 * multiple-scattering parts. */
#include <math.h>
#include <stdint.h>

#define ORC_API __attribute__((visibility("default")))

typedef struct { float x, y, z; } v3;

static const float kPi = 3.14159265358f;
static const float kGround = 6.360f, kTop = 6.460f;
enum { TW = 256, TH = 64, MW = 32, MH = 32, SW = 200, SH = 200 };

static inline v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 mul(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 muls(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 divs(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float length(v3 a) { return sqrtf(dot(a, a)); }
static inline v3 normalize(v3 a) { return muls(a, 1.0f / sqrtf(dot(a, a))); }
static inline float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
static inline float safeacos(float x) { return acosf(clampf(x, -1.0f, 1.0f)); }
static inline v3 expv(v3 a) { return mk(expf(a.x), expf(a.y), expf(a.z)); }
/* (s - s*T) / ext, componentwise: the in-segment integral of the reference */
static inline v3 seg(v3 s, v3 st, v3 ext) {
    return mk((s.x - s.x * st.x) / ext.x, (s.y - s.y * st.y) / ext.y, (s.z - s.z * st.z) / ext.z);
}

/* getMiePhase / getRayleighPhase (atmosphere.slang:25-38) */
static float mie_phase(float c) {
    const float g = 0.8f, scale = 3.0f / (8.0f * kPi);
    float num = (1.0f - g * g) * (1.0f + c * c);
    float denom = (2.0f + g * g) * powf(1.0f + g * g - 2.0f * g * c, 1.5f);
    return scale * num / denom;
}
static float rayleigh_phase(float c) { return (3.0f / (16.0f * kPi)) * (1.0f + c * c); }

/* getScatteringValues (atmosphere.slang:40-62) */
static void scattering(v3 pos, v3 *rayleigh, float *mie, v3 *extinction) {
    float alt = (length(pos) - kGround) * 1000.0f;
    float rd = expf(-alt / 8.0f), md = expf(-alt / 1.2f);
    *rayleigh = muls(mk(5.802f, 13.558f, 33.1f), rd);
    float ra = 0.0f * rd;
    *mie = 3.996f * md;
    float ma = 4.4f * md;
    v3 oz = muls(mk(0.650f, 1.881f, 0.085f), fmaxf(0.0f, 1.0f - fabsf(alt - 25.0f) / 15.0f));
    v3 r = *rayleigh;
    *extinction = add(mk(r.x + ra + *mie + ma, r.y + ra + *mie + ma, r.z + ra + *mie + ma), oz);
}

/* rayIntersectSphere (atmosphere.slang:68-77) */
static float sphere_hit(v3 ro, v3 rd, float rad) {
    float b = dot(ro, rd);
    float c = dot(ro, ro) - rad * rad;
    if (c > 0.0f && b > 0.0f) return -1.0f;
    float discr = b * b - c;
    if (discr < 0.0f) return -1.0f;
    if (discr > b * b) return -b + sqrtf(discr);
    return -b - sqrtf(discr);
}

static v3 sample_clamp(const float *lut, int w, int h, float u, float v) {
    float x = u * (float)w - 0.5f, y = v * (float)h - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    float tx = x - fx, ty = y - fy;
    int x0 = (int)fx, y0 = (int)fy;
    int xa = x0 < 0 ? 0 : (x0 > w - 1 ? w - 1 : x0), xb = x0 + 1 < 0 ? 0 : (x0 + 1 > w - 1 ? w - 1 : x0 + 1);
    int ya = y0 < 0 ? 0 : (y0 > h - 1 ? h - 1 : y0), yb = y0 + 1 < 0 ? 0 : (y0 + 1 > h - 1 ? h - 1 : y0 + 1);
    const float *a = lut + 4 * (ya * w + xa), *b = lut + 4 * (ya * w + xb);
    const float *c = lut + 4 * (yb * w + xa), *d = lut + 4 * (yb * w + xb);
    v3 r0 = add(muls(mk(a[0], a[1], a[2]), 1.0f - tx), muls(mk(b[0], b[1], b[2]), tx));
    v3 r1 = add(muls(mk(c[0], c[1], c[2]), 1.0f - tx), muls(mk(d[0], d[1], d[2]), tx));
    return add(muls(r0, 1.0f - ty), muls(r1, ty));
}

/* getValFromTLUT / getValFromMultiScattLUT (atmosphere.slang:83-113) */
static v3 lut_at(const float *lut, int w, int h, v3 pos, v3 sun) {
    float height = length(pos);
    v3 up = divs(pos, height);
    float c = dot(sun, up);
    float u = clampf(0.5f + 0.5f * c, 0.0f, 1.0f);
    float v = clampf((height - kGround) / (kTop - kGround), 0.0f, 1.0f);
    return sample_clamp(lut, w, h, u, v);
}

static inline float lerpf(float a, float b, float t) { return a + t * (b - a); }

/* getSunTransmittance (atmosphere.slang:117-140) */
static v3 sun_transmittance(v3 pos, v3 sun) {
    if (sphere_hit(pos, sun, kGround) > 0.0f) return mk(0.0f, 0.0f, 0.0f);
    float atmo = sphere_hit(pos, sun, kTop);
    float t = 0.0f;
    v3 tr = mk(1.0f, 1.0f, 1.0f);
    for (float i = 0.0f; i < 40.0f; i += 1.0f) {
        float nt = ((i + 0.3f) / 40.0f) * atmo;
        float dt = nt - t;
        t = nt;
        v3 rs, ext;
        float ms;
        scattering(add(pos, muls(sun, t)), &rs, &ms, &ext);
        tr = mul(tr, mk(expf(-dt * ext.x), expf(-dt * ext.y), expf(-dt * ext.z)));
    }
    return tr;
}

static void lut_coords(int px, int py, int w, int h, v3 *pos, v3 *sun) {
    float u = ((float)px + 0.5f) / (float)w, v = ((float)py + 0.5f) / (float)h;
    float c = 2.0f * u - 1.0f;
    float theta = safeacos(c);
    *pos = mk(0.0f, lerpf(kGround, kTop, v), 0.0f);
    *sun = normalize(mk(0.0f, c, -sinf(theta)));
}

static void put4(float *out, v3 v) { out[0] = v.x; out[1] = v.y; out[2] = v.z; out[3] = 1.0f; }

/* ---- exported ---------------------------------------------------------- */

/* rayleigh.rgb, mie, extinction.rgb at a position */
ORC_API void orc_atmo_scattering(const float *pos3, float *out7) {
    v3 rs, ext;
    float ms;
    scattering(mk(pos3[0], pos3[1], pos3[2]), &rs, &ms, &ext);
    out7[0] = rs.x; out7[1] = rs.y; out7[2] = rs.z; out7[3] = ms;
    out7[4] = ext.x; out7[5] = ext.y; out7[6] = ext.z;
}

ORC_API float orc_atmo_sphere_hit(const float *ro3, const float *rd3, float radius) {
    return sphere_hit(mk(ro3[0], ro3[1], ro3[2]), mk(rd3[0], rd3[1], rd3[2]), radius);
}

ORC_API void orc_atmo_sun_transmittance(const float *pos3, const float *dir3, float *out3) {
    v3 t = sun_transmittance(mk(pos3[0], pos3[1], pos3[2]), mk(dir3[0], dir3[1], dir3[2]));
    out3[0] = t.x; out3[1] = t.y; out3[2] = t.z;
}

/* transmittance_lut_compute (atmosphere.slang:143-166): out float[64][256][4] */
ORC_API void orc_atmo_transmittance_lut(float *out) {
    for (int py = 0; py < TH; ++py)
        for (int px = 0; px < TW; ++px) {
            v3 pos, sun;
            lut_coords(px, py, TW, TH, &pos, &sun);
            put4(out + 4 * (py * TW + px), sun_transmittance(pos, sun));
        }
}

/* getMulScattValues (atmosphere.slang:179-258) for one table texel with a
 * direction rule of sq x sq and ray marches of `steps` (the reference: 8 and
 * 20).  Larger values let the checks converge the same code to the integrals
 * it discretises. */
static v3 multiscattering_texel(const float *tlut, int px, int py, int sq, int steps) {
    const v3 ground_albedo = mk(0.3f, 0.3f, 0.3f);
    v3 pos, sun;
    lut_coords(px, py, MW, MH, &pos, &sun);
    float inv = 1.0f / (float)(sq * sq);
    const float fsteps = (float)steps;
    v3 lum_total = mk(0.0f, 0.0f, 0.0f), fms = mk(0.0f, 0.0f, 0.0f);
    for (int i = 0; i < sq; ++i)
        for (int j = 0; j < sq; ++j) {
            float theta = kPi * ((float)i + 0.5f) / (float)sq;
            float phi = safeacos(1.0f - 2.0f * ((float)j + 0.5f) / (float)sq);
            float cp = cosf(phi), sp = sinf(phi), ct = cosf(theta), st = sinf(theta);
            v3 dir = mk(sp * st, cp, sp * ct);
            float atmo = sphere_hit(pos, dir, kTop), grd = sphere_hit(pos, dir, kGround);
            float tmax = grd > 0.0f ? grd : atmo;
            float c = dot(dir, sun);
            float mp = mie_phase(c), rp = rayleigh_phase(-c);
            v3 lum = mk(0.0f, 0.0f, 0.0f), lum_factor = mk(0.0f, 0.0f, 0.0f), tr = mk(1.0f, 1.0f, 1.0f);
            float t = 0.0f;
            for (float s = 0.0f; s < fsteps; s += 1.0f) {
                float nt = ((s + 0.3f) / fsteps) * tmax;
                float dt = nt - t;
                t = nt;
                v3 np = add(pos, muls(dir, t));
                v3 rs, ext;
                float ms;
                scattering(np, &rs, &ms, &ext);
                v3 stt = expv(mk(-dt * ext.x, -dt * ext.y, -dt * ext.z));
                v3 snp = mk(rs.x + ms, rs.y + ms, rs.z + ms);
                lum_factor = add(lum_factor, mul(tr, seg(snp, stt, ext)));
                v3 sunt = lut_at(tlut, TW, TH, np, sun);
                v3 ris = muls(rs, rp);
                float mis = ms * mp;
                v3 ins = mul(mk(ris.x + mis, ris.y + mis, ris.z + mis), sunt);
                lum = add(lum, mul(seg(ins, stt, ext), tr));
                tr = mul(tr, stt);
            }
            if (grd > 0.0f) {
                v3 hit = add(pos, muls(dir, grd));
                if (dot(pos, sun) > 0.0f) {
                    hit = muls(normalize(hit), kGround);
                    lum = add(lum, mul(mul(tr, ground_albedo), lut_at(tlut, TW, TH, hit, sun)));
                }
            }
            fms = add(fms, muls(lum_factor, inv));
            lum_total = add(lum_total, muls(lum, inv));
        }
    return mk(lum_total.x / (1.0f - fms.x), lum_total.y / (1.0f - fms.y), lum_total.z / (1.0f - fms.z));
}

/* multiscatt_lut_compute (atmosphere.slang:172-258): out float[32][32][4] */
ORC_API void orc_atmo_multiscattering_lut(const float *tlut, float *out) {
    for (int py = 0; py < MH; ++py)
        for (int px = 0; px < MW; ++px) put4(out + 4 * (py * MW + px), multiscattering_texel(tlut, px, py, 8, 20));
}

/* One texel of the table with another quadrature resolution (for the checks). */
ORC_API void orc_atmo_multiscattering_texel(const float *tlut, int px, int py, int sq, int steps, float *out3) {
    v3 v = multiscattering_texel(tlut, px, py, sq, steps);
    out3[0] = v.x;
    out3[1] = v.y;
    out3[2] = v.z;
}

/* sky_view_lut_compute + raymarchScattering (atmosphere.slang:267-397):
 * out float[200][200][4] */
ORC_API void orc_atmo_sky_view_lut(const float *tlut, const float *mslut, const float *sun_direction3, float *out) {
    v3 sd = mk(sun_direction3[0], sun_direction3[1], sun_direction3[2]);
    for (int py = 0; py < SH; ++py)
        for (int px = 0; px < SW; ++px) {
            float u = ((float)px + 0.5f) / (float)SW, v = ((float)py + 0.5f) / (float)SH;
            float az = (u - 0.5f) * 2.0f * kPi;
            float adj;
            if (v < 0.5f) {
                float c = 1.0f - 2.0f * v;
                adj = -c * c;
            } else {
                float c = v * 2.0f - 1.0f;
                adj = c * c;
            }
            v3 view = mk(0.0f, kGround + 0.0002f, 0.0f);
            float height = length(view);
            v3 up = divs(view, height);
            float horizon = safeacos(sqrtf(height * height - kGround * kGround) / height) - 0.5f * kPi;
            float alt = adj * 0.5f * kPi - horizon;
            float ca = cosf(alt);
            v3 dir = mk(ca * sinf(az), sinf(alt), -ca * cosf(az));
            float sun_alt = (0.5f * kPi) - acosf(dot(sd, up));
            v3 sun = mk(0.0f, sinf(sun_alt), -cosf(sun_alt));
            float atmo = sphere_hit(view, dir, kTop), grd = sphere_hit(view, dir, kGround);
            float tmax = grd < 0.0f ? atmo : grd;
            float c = dot(dir, sun);
            float mp = mie_phase(c), rp = rayleigh_phase(-c);
            v3 lum = mk(0.0f, 0.0f, 0.0f), tr = mk(1.0f, 1.0f, 1.0f);
            float t = 0.0f;
            for (float i = 0.0f; i < 32.0f; i += 1.0f) {
                float nt = ((i + 0.3f) / 32.0f) * tmax;
                float dt = nt - t;
                t = nt;
                v3 np = add(view, muls(dir, t));
                v3 rs, ext;
                float ms;
                scattering(np, &rs, &ms, &ext);
                v3 stt = expv(mk(-dt * ext.x, -dt * ext.y, -dt * ext.z));
                v3 sunt = lut_at(tlut, TW, TH, np, sun);
                v3 psi = lut_at(mslut, MW, MH, np, sun);
                v3 ris = mul(rs, add(muls(sunt, rp), psi));
                v3 mis = muls(add(muls(sunt, mp), psi), ms);
                v3 ins = add(ris, mis);
                lum = add(lum, mul(seg(ins, stt, ext), tr));
                tr = mul(tr, stt);
            }
            put4(out + 4 * (py * SW + px), lum);
        }
}

/* get_sun_transmittance_at_view (scene.slang:997-1009) */
ORC_API void orc_atmo_view_transmittance(const float *tlut, const float *sun_direction3, float *out3) {
    v3 view = mk(0.0f, kGround + 0.0002f, 0.0f);
    v3 t = lut_at(tlut, TW, TH, view, normalize(mk(sun_direction3[0], sun_direction3[1], sun_direction3[2])));
    out3[0] = t.x; out3[1] = t.y; out3[2] = t.z;
}
