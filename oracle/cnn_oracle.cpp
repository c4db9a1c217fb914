// ORACLE — test infrastructure only (tests/, smoke() and bench.py's CPU
// baseline).  Never linked into or called by the product path.
//
// CPU restatement of the near-separatrix classifier (msceer_cnn_*): the
// windowed normalisation, the three 3x3x3 convolutions, the sigmoid head, the
// weighted binary cross-entropy and its gradient, all in FP64.  With
// quant = 1 it rounds to BF16 exactly where the GPU kernels store BF16
// (convolution weights, activations after ReLU, back-propagated gradients
// before each transposed convolution), so the GPU must agree up to FP32
// accumulation order; with quant = 0 nothing is rounded, which is what the
// finite-difference tests check.  This network has no counterpart in the
// reference; its definition is the kernels' own (msceer_amd/csrc/cnn.hip).
//
// Parameter layout: W1[27][16] b1[16] W2[27][32][16] b2[32] W3[27][16][32]
// b3[16] wh[16] bh; tap t = (dx+1) + 3(dy+1) + 9(dz+1) reads (x+dx, y+dy, z+dz).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int kW1 = 27 * 16, kW2 = 27 * 32 * 16, kW3 = 27 * 16 * 32;
constexpr int kOffW1 = 0, kOffB1 = kW1, kOffW2 = kOffB1 + 16, kOffB2 = kOffW2 + kW2;
constexpr int kOffW3 = kOffB2 + 32, kOffB3 = kOffW3 + kW3, kOffWh = kOffB3 + 16;
constexpr int kNumParams = kOffWh + 17;

// round to the nearest BF16, ties to even (as the GPU's conversion)
double bf(double x, int quant) {
    if (!quant) return x;
    float f = (float)x;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) != 0x7F800000u) u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    std::memcpy(&f, &u, 4);
    return (double)f;
}

struct Vol {
    long long X, Y, Z;
    long long n() const { return X * Y * Z; }
    bool in(long long x, long long y, long long z) const {
        return x >= 0 && x < X && y >= 0 && y < Y && z >= 0 && z < Z;
    }
    long long id(long long x, long long y, long long z) const { return x + X * (y + Y * z); }
};

void normalize(const float* f, const Vol& g, std::vector<double>& out) {
    out.assign((size_t)g.n(), 0.0);
    for (long long z = 0; z < g.Z; z++)
        for (long long y = 0; y < g.Y; y++)
            for (long long x = 0; x < g.X; x++) {
                double s1 = 0.0, s2 = 0.0, c = 0.0;
                for (long long dz = -4; dz <= 4; dz++)
                    for (long long dy = -4; dy <= 4; dy++)
                        for (long long dx = -4; dx <= 4; dx++) {
                            if (!g.in(x + dx, y + dy, z + dz)) continue;
                            const double a = f[g.id(x + dx, y + dy, z + dz)];
                            s1 += a;
                            s2 += a * a;
                            c += 1.0;
                        }
                const double mean = s1 / c;
                double var = s2 / c - mean * mean;
                if (var < 0.0) var = 0.0;
                out[(size_t)g.id(x, y, z)] = (double)(float)(((double)f[g.id(x, y, z)] - mean) / (std::sqrt(var) + 1e-6));
            }
}

// out[v][co] = sum_t sum_ci w(t, co, ci) in[v + off_t][ci]  (+ bias, relu)
template <typename W>
void conv(const std::vector<double>& in, int cin, const Vol& g, int cout, W w, std::vector<double>& out, int sign) {
    out.assign((size_t)g.n() * cout, 0.0);
    for (long long z = 0; z < g.Z; z++)
        for (long long y = 0; y < g.Y; y++)
            for (long long x = 0; x < g.X; x++) {
                double* o = &out[(size_t)g.id(x, y, z) * cout];
                for (int t = 0; t < 27; t++) {
                    const long long xx = x + sign * (t % 3 - 1), yy = y + sign * ((t / 3) % 3 - 1), zz = z + sign * (t / 9 - 1);
                    if (!g.in(xx, yy, zz)) continue;
                    const double* a = &in[(size_t)g.id(xx, yy, zz) * cin];
                    for (int co = 0; co < cout; co++)
                        for (int ci = 0; ci < cin; ci++) o[co] += w(t, co, ci) * a[ci];
                }
            }
}

struct Net {
    const float* p;
    int quant;
    std::vector<double> w1, w2, w3;
    Net(const float* params, int q) : p(params), quant(q) {
        w1.resize(kW1);
        w2.resize(kW2);
        w3.resize(kW3);
        for (int i = 0; i < kW1; i++) w1[i] = bf(p[kOffW1 + i], q);
        for (int i = 0; i < kW2; i++) w2[i] = bf(p[kOffW2 + i], q);
        for (int i = 0; i < kW3; i++) w3[i] = bf(p[kOffW3 + i], q);
    }
};

struct Acts {
    std::vector<double> n, a1, a2, a3, logit;
};

void forward(const Net& net, const float* f, const Vol& g, Acts& A) {
    const int q = net.quant;
    normalize(f, g, A.n);
    conv(A.n, 1, g, 16, [&](int t, int co, int) { return net.w1[t * 16 + co]; }, A.a1, 1);
    for (size_t i = 0; i < A.a1.size(); i++) {
        const double z = A.a1[i] + net.p[kOffB1 + (int)(i % 16)];
        A.a1[i] = bf(z > 0 ? z : 0.0, q);
    }
    conv(A.a1, 16, g, 32, [&](int t, int co, int ci) { return net.w2[(t * 32 + co) * 16 + ci]; }, A.a2, 1);
    for (size_t i = 0; i < A.a2.size(); i++) {
        const double z = A.a2[i] + net.p[kOffB2 + (int)(i % 32)];
        A.a2[i] = bf(z > 0 ? z : 0.0, q);
    }
    conv(A.a2, 32, g, 16, [&](int t, int co, int ci) { return net.w3[(t * 16 + co) * 32 + ci]; }, A.a3, 1);
    for (size_t i = 0; i < A.a3.size(); i++) {
        const double z = A.a3[i] + net.p[kOffB3 + (int)(i % 16)];
        A.a3[i] = bf(z > 0 ? z : 0.0, q);
    }
    A.logit.assign((size_t)g.n(), 0.0);
    for (long long v = 0; v < g.n(); v++) {
        double s = net.p[kOffWh + 16];
        for (int c = 0; c < 16; c++) s += (double)net.p[kOffWh + c] * A.a3[(size_t)v * 16 + c];
        A.logit[(size_t)v] = s;
    }
}

}  // namespace

extern "C" {

int oracle_cnn_num_params(void) { return kNumParams; }

int oracle_cnn_normalize(const float* f, long long X, long long Y, long long Z, float* out) {
    std::vector<double> n;
    normalize(f, Vol{X, Y, Z}, n);
    for (size_t i = 0; i < n.size(); i++) out[i] = (float)n[i];
    return 0;
}

int oracle_cnn_predict(const float* params, const float* f, long long X, long long Y, long long Z, int quant,
                       double* prob) {
    const Vol g{X, Y, Z};
    Net net(params, quant);
    Acts A;
    forward(net, f, g, A);
    for (long long v = 0; v < g.n(); v++) prob[v] = 1.0 / (1.0 + std::exp(-A.logit[(size_t)v]));
    return 0;
}

// Mean weighted BCE and its gradient (kNumParams doubles).
int oracle_cnn_gradient(const float* params, const float* f, long long X, long long Y, long long Z,
                        const unsigned char* target, double pos_w, int quant, double* loss, double* grad) {
    const Vol g{X, Y, Z};
    const long long N = g.n();
    Net net(params, quant);
    Acts A;
    forward(net, f, g, A);
    const int q = quant;
    for (int i = 0; i < kNumParams; i++) grad[i] = 0.0;
    double L = 0.0;
    std::vector<double> dz3((size_t)N * 16, 0.0);
    for (long long v = 0; v < N; v++) {
        const double z = A.logit[(size_t)v];
        const double t = target[v] ? 1.0 : 0.0, w = target[v] ? pos_w : 1.0;
        const double lp = z >= 0 ? -std::log1p(std::exp(-z)) : z - std::log1p(std::exp(z));
        const double ln = lp - z;
        L += -w * (t * lp + (1.0 - t) * ln);
        const double p = 1.0 / (1.0 + std::exp(-z));
        const double dl = w * (p - t) / (double)N;
        grad[kOffWh + 16] += dl;
        for (int c = 0; c < 16; c++) {
            const double a = A.a3[(size_t)v * 16 + c];
            grad[kOffWh + c] += dl * a;
            dz3[(size_t)v * 16 + c] = a > 0 ? bf(dl * params[kOffWh + c], q) : 0.0;
        }
    }
    *loss = L / (double)N;
    // weight gradient of a convolution: dW(t, co, ci) += in[v + off_t][ci] dz[v][co]
    auto wgrad = [&](const std::vector<double>& in, int cin, const std::vector<double>& dz, int cout, int off_w,
                     int off_b) {
        for (long long zz = 0; zz < g.Z; zz++)
            for (long long yy = 0; yy < g.Y; yy++)
                for (long long xx = 0; xx < g.X; xx++) {
                    const long long v = g.id(xx, yy, zz);
                    for (int co = 0; co < cout; co++) {
                        const double d = dz[(size_t)v * cout + co];
                        if (d == 0.0) continue;
                        grad[off_b + co] += d;
                        for (int t = 0; t < 27; t++) {
                            const long long x2 = xx + t % 3 - 1, y2 = yy + (t / 3) % 3 - 1, z2 = zz + t / 9 - 1;
                            if (!g.in(x2, y2, z2)) continue;
                            const double* a = &in[(size_t)g.id(x2, y2, z2) * cin];
                            for (int ci = 0; ci < cin; ci++) grad[off_w + (t * cout + co) * cin + ci] += d * a[ci];
                        }
                    }
                }
    };
    wgrad(A.a2, 32, dz3, 16, kOffW3, kOffB3);
    // dz2 = (transposed conv of dz3 with W3) gated by a2 > 0
    std::vector<double> da2, dz2;
    conv(dz3, 16, g, 32, [&](int t, int co2, int ci2) { return net.w3[(t * 16 + ci2) * 32 + co2]; }, da2, -1);
    dz2.resize(da2.size());
    for (size_t i = 0; i < da2.size(); i++) dz2[i] = A.a2[i] > 0 ? bf(da2[i], q) : 0.0;
    wgrad(A.a1, 16, dz2, 32, kOffW2, kOffB2);
    std::vector<double> da1, dz1;
    conv(dz2, 32, g, 16, [&](int t, int co1, int ci1) { return net.w2[(t * 32 + ci1) * 16 + co1]; }, da1, -1);
    dz1.resize(da1.size());
    for (size_t i = 0; i < da1.size(); i++) dz1[i] = A.a1[i] > 0 ? bf(da1[i], q) : 0.0;
    wgrad(A.n, 1, dz1, 16, kOffW1, kOffB1);
    return 0;
}

}  // extern "C"
