// CPU emulation of the mixed dense walk with each stage's arithmetic chosen
// separately, for an error budget against FP64.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../dbgenzpaper_amd/csrc/tailq.h"

using namespace dbg;

struct Cfg {
    int pt32;       // 32-bit fixed-point points, FP32 w
    int pbits;      // panel operands: 0 exact, 22 split16 (two FP16), 33 three FP16, 24 FP32
    int pacc;       // panel accumulation: 0 exact, 1 FP32 running per MFMA, 2 FP32 per group, exact across
    int pgroup;     // chunks (of 32 k) per FP32 group for pacc = 2
    int inpanel;    // in-panel update: 0 exact, 1 FP32 ops + FP32 acc, 2 FP32 ops exact acc
    int inblock;    // in-block dot: 0 exact, 1 FP32 fma chain, 2 FP32 ops exact acc
    int z;          // 0 exact, 1 FP32 (b, inv rounded; s rounded), 2 exact z rounded to FP32, 3 z with first-order mass correction
    int mass;       // 0 FP64 erfc, 1 tail_mass_f32
    int quant;      // 0 FP64, 1 tail_quantile_f32
    int weight;     // 0 FP64 log sum, 1 FP32 mantissa product
    int state;      // 0 FP64 state, 1 FP32 scaled state
    int ybits;      // y values stored for later use (contraction) rounded: 0 none, else bits
};

static inline double f32(double x) { return (double)(float)x; }
// round to FP16 (11 significant bits, subnormal quantum 2^-24), nearest-even
static inline double h16(double x)
{
    if (x == 0 || !isfinite(x)) return x;
    double ax = fabs(x);
    if (ax > 65504.0) return copysign(65504.0, x);
    int e = 0;
    frexp(ax, &e);
    int qe = e - 11;
    if (qe < -24) qe = -24;
    const double qv = ldexp(1.0, qe);
    return copysign(nearbyint(ax / qv) * qv, x);
}
static inline double rbits(double x, int B)
{
    if (x == 0 || !isfinite(x)) return x;
    int e = 0;
    frexp(x, &e);
    const double qv = ldexp(1.0, e - B);
    return nearbyint(x / qv) * qv;
}
// operand as the sum of its FP16 parts (prescaled value v)
static inline double split_parts(double v, int nparts, double* parts)
{
    double r = v, s = 0;
    for (int i = 0; i < nparts; ++i) {
        parts[i] = h16(r);
        r -= parts[i];
        s += parts[i];
    }
    return s;
}

static double quantile64(double t)
{
    if (t <= 0) return -INFINITY;
    int e = 0;
    double mt = frexp(t, &e);
    float y0 = tail_quantile_f32((float)mt, -e);
    long double y = y0;
    long double lt = logl((long double)t);
    for (int it = 0; it < 4; ++it) {
        long double P = 0.5L * erfcl(-y / sqrtl(2.0L));
        long double ph = expl(-0.5L * y * y) / sqrtl(2.0L * 3.14159265358979323846L);
        long double f = logl(P) - lt;
        y -= f * P / ph;
    }
    return (double)y;
}

struct St {
    double ea, eb, dc;
    int mode, k;
};

// mass of one bound: m * 2^-k = erfc(z)/2, z >= 0; zl: low part of z (first-order correction)
static double mass_one(double z, double zl, const Cfg& c, int& k)
{
    if (c.mass) {
        float m = tail_mass_f32((float)z, k);
        if (c.z == 3) {
            // d log erfc(z) / dz = -2 e^{-z^2} / (sqrt(pi) erfc(z))
            double der = -2.0 * exp(-z * z) / (sqrt(M_PI) * erfc(z));
            return (double)m * (1.0 + der * zl);
        }
        return m;
    }
    k = 0;
    return 0.5 * erfc(z + zl);
}

static void masses(double za, double zb, double zal, double zbl, const Cfg& c, St& st)
{
    int ka = 0, kb = 0;
    double ma = isinf(za) ? 0.0 : mass_one(fabs(za), za < 0 ? -zal : zal, c, ka);
    double mb = isinf(zb) ? 0.0 : mass_one(fabs(zb), zb < 0 ? -zbl : zbl, c, kb);
    const int mode = (zb <= 0.0) ? 0 : (za >= 0.0) ? 1 : 2;
    double ea = ldexp(ma, -ka), eb = ldexp(mb, -kb);
    double dc = (mode == 0) ? eb - ea : (mode == 1) ? ea - eb : (1.0 - ea) - eb;
    if (c.state) {
        double mx = fmax(fmax(ea, eb), fabs(dc));
        int k = 0;
        if (mx > 0.0 && mx < 1e-30) { int e = 0; frexp(mx, &e); k = -e; }
        st.ea = f32(ldexp(ea, k)); st.eb = f32(ldexp(eb, k)); st.dc = f32(ldexp(dc, k)); st.k = k;
        if (c.mass) {  // FP32 arithmetic of the masses in the scaled state
            float fa = (float)ldexp(ea, k), fb = (float)ldexp(eb, k);
            st.dc = (mode == 0) ? fb - fa : (mode == 1) ? fa - fb : (1.0f - fa) - fb;
        }
    } else {
        st.ea = ea; st.eb = eb; st.dc = dc; st.k = 0;
    }
    st.mode = mode;
}

static double sample_y(const St& st, double w, double wc, const Cfg& c)
{
    if (c.quant) {
        float u_lo = fmaf((float)w, (float)st.dc, (float)st.ea);
        float u_hi = fmaf((float)wc, (float)st.dc, (float)st.eb);
        bool upper = (st.mode == 1) || (st.mode == 2 && u_hi < u_lo);
        float t = fmaxf(upper ? u_hi : u_lo, 1.17549435e-38f);
        float y = tail_quantile_f32(t, st.k);
        return upper ? -y : y;
    }
    double u_lo = fma(w, st.dc, st.ea);
    double u_hi = fma(wc, st.dc, st.eb);
    bool upper = (st.mode == 1) || (st.mode == 2 && u_hi < u_lo);
    double t = ldexp(upper ? u_hi : u_lo, -st.k);
    double y = quantile64(t);
    return upper ? -y : y;
}

extern "C" int budget_walk(int n, const double* U, const double* a, const double* b, const double* q,
                           const double* shift, int64_t m, const Cfg* cfg, double* logw)
{
    const Cfg c = *cfg;
    const int nd = n - 1;
    const int PW = 128, SB = 16, KC = 32;
    std::vector<uint64_t> q64(nd);
    std::vector<uint32_t> s32(nd);
    for (int g = 0; g < nd; ++g) {
        q64[g] = (uint64_t)ldexp(q[g] - floor(q[g]), 64);
        s32[g] = (uint32_t)(uint64_t)ldexp(shift[g] - floor(shift[g]), 32);
    }
    std::vector<double> inv(n);
    for (int i = 0; i < n; ++i) inv[i] = 1.0 / (U[(size_t)i * n + i] * 1.4142135623730951);
    const int nparts_u = c.pbits == 22 ? 2 : c.pbits == 33 ? 3 : 0;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t jj = 0; jj < m; ++jj) {
        const int64_t j = jj + 1;
        std::vector<double> y(nd), ys(nd), acc(n, 0.0);
        St st;
        {
            Cfg c0 = c;
            c0.mass = 0;
            c0.z = 0;
            double za = a[0] * inv[0], zb = b[0] * inv[0];
            masses(za, zb, 0, 0, c0, st);
        }
        double lw;
        float pm = 0.f;
        int pe = 0;
        {
            double dc = ldexp(st.dc, -st.k);
            lw = log(dc);
            int e = 0;
            pm = frexpf((float)st.dc, &e);
            pe = e - st.k;
        }
        for (int g0 = 0; g0 < nd; g0 += PW) {
            const int t1 = std::min(n, g0 + PW + 1);
            for (int t = g0 + 1; t < t1; ++t) {
                double s = 0.0;
                float accf = 0.f;
                double grp = 0.0;
                float grpf = 0.f;
                for (int ch = 0; ch < g0 / KC; ++ch) {
                    double part[9] = {0};
                    int np = 0;
                    for (int e = 0; e < KC; ++e) {
                        const int l = ch * KC + e;
                        const double yv = ys[l], uv = U[(size_t)t * n + l];
                        if (nparts_u) {
                            double yp[3], up[3];
                            split_parts(yv * 64.0, nparts_u, yp);
                            split_parts(uv * 8192.0, nparts_u, up);
                            if (nparts_u == 2) {
                                part[0] += yp[0] * up[0]; part[1] += yp[0] * up[1]; part[2] += yp[1] * up[0];
                                np = 3;
                            } else {  // products with i + j <= 2
                                part[0] += yp[0] * up[0]; part[1] += yp[0] * up[1]; part[2] += yp[1] * up[0];
                                part[3] += yp[0] * up[2]; part[4] += yp[1] * up[1]; part[5] += yp[2] * up[0];
                                np = 6;
                            }
                        } else if (c.pbits == 24) {
                            part[0] += f32(yv * 64.0) * f32(uv * 8192.0);
                            np = 1;
                        } else {
                            part[0] += yv * uv * 524288.0;
                            np = 1;
                        }
                    }
                    if (c.pacc == 1) {
                        for (int r = 0; r < np; ++r) accf = (float)((double)accf + part[r]);
                    } else if (c.pacc == 2) {
                        for (int r = 0; r < np; ++r) grpf = (float)((double)grpf + part[r]);
                        if ((ch + 1) % c.pgroup == 0) { grp += grpf; grpf = 0.f; }
                    } else {
                        for (int r = 0; r < np; ++r) s += part[r];
                    }
                }
                if (c.pacc == 1) s = accf;
                else if (c.pacc == 2) s = grp + grpf;
                acc[t] = s * (1.0 / 524288.0);
            }
            for (int gb = g0; gb < std::min(nd, g0 + PW); gb += SB) {
                const int cnt = std::min(SB, nd - gb);
                for (int cc = 0; cc < cnt; ++cc) {
                    const int g = gb + cc, t = g + 1;
                    double w, wc;
                    if (c.pt32) {
                        const uint64_t nj = (uint64_t)j;
                        const uint32_t nlo = (uint32_t)nj, nhi = (uint32_t)(nj >> 32);
                        const uint32_t qlo = (uint32_t)q64[g], qhi = (uint32_t)(q64[g] >> 32);
                        const uint32_t x = (uint32_t)(((uint64_t)nlo * qlo) >> 32) + nlo * qhi + nhi * qlo;
                        const uint32_t u = x + s32[g];
                        const uint32_t tw = (u & 0x80000000u) ? (u << 1) : (0u - (u << 1));
                        const float kS = 2.3283064365386963e-10f;
                        w = (u == 0u) ? 1.0f : (float)tw * kS;
                        wc = (u == 0u) ? 0.0f : (tw == 0u ? 1.0f : (float)(0u - tw) * kS);
                    } else {
                        double x = (double)j * q[g];
                        x = x - floor(x);
                        double u = shift[g] + x;
                        if (u > 1.0) u -= 1.0;
                        w = fabs(fma(u, 2.0, -1.0));
                        wc = 1.0 - w;
                    }
                    y[g] = sample_y(st, w, wc, c);
                    ys[g] = c.ybits ? rbits(y[g], c.ybits) : y[g];
                    double s = acc[t];
                    if (c.inblock == 1) {
                        float sf = (float)s;
                        for (int l = gb; l <= g; ++l) sf = fmaf((float)ys[l], (float)U[(size_t)t * n + l], sf);
                        s = sf;
                    } else if (c.inblock == 2) {
                        for (int l = gb; l <= g; ++l) s += f32(ys[l]) * f32(U[(size_t)t * n + l]);
                    } else {
                        for (int l = gb; l <= g; ++l) s += U[(size_t)t * n + l] * ys[l];
                    }
                    if (t < n) {
                        double za, zb, zal = 0, zbl = 0;
                        if (c.z == 1) {
                            za = (double)(((float)a[t] - (float)s) * (float)inv[t]);
                            zb = (double)(((float)b[t] - (float)s) * (float)inv[t]);
                        } else if (c.z == 2 || c.z == 3) {
                            double ze = (a[t] - s) * inv[t], zeb = (b[t] - s) * inv[t];
                            za = f32(ze); zb = f32(zeb);
                            if (c.z == 3) { zal = ze - za; zbl = zeb - zb; }
                        } else {
                            za = (a[t] - s) * inv[t];
                            zb = (b[t] - s) * inv[t];
                        }
                        masses(za, zb, zal, zbl, c, st);
                        double dc = ldexp(st.dc, -st.k);
                        lw += log(dc);
                        int e = 0;
                        pm = frexpf(pm * (float)st.dc, &e);
                        pe += e - st.k;
                    }
                }
                const int tend = std::min(n, g0 + PW + 1);
                for (int t = gb + cnt + 1; t < tend; ++t) {
                    if (c.inpanel == 1) {
                        float accf = (float)acc[t];
                        for (int k4 = 0; k4 < cnt; k4 += 4) {
                            double part = 0.0;
                            for (int e = k4; e < std::min(cnt, k4 + 4); ++e)
                                part += f32(ys[gb + e]) * f32(U[(size_t)t * n + gb + e]);
                            accf = (float)((double)accf + part);
                        }
                        acc[t] = accf;
                    } else if (c.inpanel == 2) {
                        double s = acc[t];
                        for (int e = 0; e < cnt; ++e) s += f32(ys[gb + e]) * f32(U[(size_t)t * n + gb + e]);
                        acc[t] = s;
                    } else {
                        double s = acc[t];
                        for (int e = 0; e < cnt; ++e) s += U[(size_t)t * n + gb + e] * ys[gb + e];
                        acc[t] = s;
                    }
                }
            }
        }
        logw[jj] = c.weight ? (log((double)pm) + pe * 0.69314718055994530942) : lw;
    }
    return 0;
}
