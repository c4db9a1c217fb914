// Sanitizer driver (test infrastructure): the CPU oracle and the host
// emulation of the lower-star kernel built with AddressSanitizer and
// UndefinedBehaviorSanitizer, run over small volumes with boundaries, ties
// and a budget-bound streamline, so that an out-of-bounds access or undefined
// arithmetic in the checkers themselves fails loudly.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

extern "C" {
int oracle_steepest_grad(const float* f, long long X, long long Y, long long Z, unsigned char* grad);
int oracle_ondemand_asc(const float* f, long long X, long long Y, long long Z, float ethr, float gthr, int max_iters,
                        float presimp, int mode, unsigned char* grad, int* labels, signed char* restriction,
                        long long* stats);
int oracle_cnn_num_params(void);
int oracle_cnn_gradient(const float* params, const float* f, long long X, long long Y, long long Z,
                        const unsigned char* target, double pos_w, int quant, double* loss, double* grad);
int emul_lstar_prepass(const float* f, long long X, long long Y, long long Z, const int8_t* restriction,
                       uint8_t* grad);
int emul_check_product(const float* f, long long X, long long Y, long long Z, long long* counts);
}

static std::vector<float> field(long long X, long long Y, long long Z, int kind) {
    std::vector<float> f((size_t)(X * Y * Z));
    uint64_t s = 0x9E3779B97F4A7C15ull * (uint64_t)(kind + 1);
    for (long long i = 0; i < X * Y * Z; i++) {
        s ^= s << 13; s ^= s >> 7; s ^= s << 17;
        const long long x = i % X, y = (i / X) % Y, z = i / (X * Y);
        const double smooth = std::sin(0.4 * x) * std::cos(0.3 * y) + 0.2 * std::sin(0.5 * z);
        const double noise = (double)(s >> 11) / 9007199254740992.0 - 0.5;
        f[(size_t)i] = kind == 0 ? (float)(smooth + 0.05 * noise) : (float)std::floor(4.0 * noise);  // kind 1: ties
    }
    return f;
}

int main() {
    const long long dims[][3] = {{9, 7, 5}, {2, 3, 4}, {17, 11, 6}};
    for (const auto& d : dims)
        for (int kind = 0; kind < 2; kind++) {
            const long long X = d[0], Y = d[1], Z = d[2];
            const size_t C = (size_t)((2 * X - 1) * (2 * Y - 1) * (2 * Z - 1));
            std::vector<float> f = field(X, Y, Z, kind);
            std::vector<unsigned char> g(C), g2(C);
            std::vector<int> lab((size_t)(X * Y * Z));
            std::vector<signed char> r(C);
            long long st[8] = {0}, cnt[2] = {0};
            if (oracle_steepest_grad(f.data(), X, Y, Z, g.data())) return 1;
            if (emul_lstar_prepass(f.data(), X, Y, Z, nullptr, g2.data())) return 1;
            if (emul_check_product(f.data(), X, Y, Z, cnt) || cnt[1] != 0) return 1;
            for (int mode = 0; mode < 3; mode++)
                if (oracle_ondemand_asc(f.data(), X, Y, Z, 1e-3f, 1e-6f, mode == 1 ? 8 : 4096, 0.0f, mode, g.data(),
                                        lab.data(), r.data(), st))
                    return 1;
            if (emul_lstar_prepass(f.data(), X, Y, Z, r.data(), g2.data())) return 1;
        }
    const long long X = 5, Y = 4, Z = 3;
    std::vector<float> f = field(X, Y, Z, 0), p((size_t)oracle_cnn_num_params(), 0.01f);
    std::vector<unsigned char> t((size_t)(X * Y * Z), 1);
    std::vector<double> grad(p.size());
    double loss = 0.0;
    if (oracle_cnn_gradient(p.data(), f.data(), X, Y, Z, t.data(), 2.0, 1, &loss, grad.data())) return 1;
    std::printf("sanitized oracle and emulation: ok\n");
    return 0;
}
