// Near-separatrix classifier: a small 3-D CNN with hand-written kernels.
//
// Input: the value field normalised over the 9x9x9 window centred on each
// vertex (window clamped to the domain): n = (f - mean) / (std + 1e-6).
// Layers (3x3x3 taps, zero padding, channels-last activations in BF16):
//   conv1  1 -> 16, ReLU        (VALU, FP32)
//   conv2 16 -> 32, ReLU        (MFMA 32x32x16 bf16, FP32 accumulation)
//   conv3 32 -> 16, ReLU        (MFMA 16x16x32 bf16, FP32 accumulation)
//   head  16 -> 1, sigmoid      (fused into conv3 for inference)
// Weights are kept in FP32 (Adam state) and packed to BF16 for the
// convolutions; biases and the head stay FP32.  Tap t = (dx+1) + 3(dy+1) +
// 9(dz+1) reads the input at (x+dx, y+dy, z+dz).
//
// Training: weighted binary cross-entropy (mean over voxels), backward by
// transposed convolutions on the same MFMA kernels (taps mirrored, weights
// transposed), weight gradients on VALU with per-block partial sums reduced
// in a fixed order (deterministic), Adam.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "cnn.hpp"
#include "ctx.hpp"

namespace msceer {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kT = 256;

__device__ __forceinline__ uint16_t f2bf(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)(u >> 16);  // inf / nan
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

struct Geo {
    int X, Y, Z;
    int64_t sxy;
    __host__ __device__ __forceinline__ int64_t verts() const { return sxy * Z; }
};

__device__ __forceinline__ int tap_dx(int t) { return t % 3 - 1; }
__device__ __forceinline__ int tap_dy(int t) { return (t / 3) % 3 - 1; }
__device__ __forceinline__ int tap_dz(int t) { return t / 9 - 1; }

// ---- local normalisation (separable 9-wide box sums in FP64) ----------------

__global__ void __launch_bounds__(kT) norm_x_kernel(const float* __restrict__ f, Geo g, double2* __restrict__ out) {
    const int64_t n = g.verts();
    for (int64_t v = (int64_t)blockIdx.x * kT + threadIdx.x; v < n; v += (int64_t)gridDim.x * kT) {
        const int x = (int)(v % g.X);
        double s1 = 0.0, s2 = 0.0;
        for (int d = -4; d <= 4; d++) {
            const int xx = x + d;
            if (xx < 0 || xx >= g.X) continue;
            const double a = (double)f[v + d];
            s1 += a;
            s2 += a * a;
        }
        out[v] = make_double2(s1, s2);
    }
}

__global__ void __launch_bounds__(kT) norm_y_kernel(const double2* __restrict__ in, Geo g, double2* __restrict__ out) {
    const int64_t n = g.verts();
    for (int64_t v = (int64_t)blockIdx.x * kT + threadIdx.x; v < n; v += (int64_t)gridDim.x * kT) {
        const int y = (int)((v / g.X) % g.Y);
        double s1 = 0.0, s2 = 0.0;
        for (int d = -4; d <= 4; d++) {
            const int yy = y + d;
            if (yy < 0 || yy >= g.Y) continue;
            const double2 a = in[v + (int64_t)d * g.X];
            s1 += a.x;
            s2 += a.y;
        }
        out[v] = make_double2(s1, s2);
    }
}

__device__ __forceinline__ int win_count(int c, int n) {
    const int lo = c - 4 < 0 ? 0 : c - 4, hi = c + 4 >= n ? n - 1 : c + 4;
    return hi - lo + 1;
}

__global__ void __launch_bounds__(kT)
norm_z_kernel(const double2* __restrict__ in, const float* __restrict__ f, Geo g, float* __restrict__ out) {
    const int64_t n = g.verts();
    for (int64_t v = (int64_t)blockIdx.x * kT + threadIdx.x; v < n; v += (int64_t)gridDim.x * kT) {
        const int x = (int)(v % g.X), y = (int)((v / g.X) % g.Y), z = (int)(v / g.sxy);
        double s1 = 0.0, s2 = 0.0;
        for (int d = -4; d <= 4; d++) {
            const int zz = z + d;
            if (zz < 0 || zz >= g.Z) continue;
            const double2 a = in[v + (int64_t)d * g.sxy];
            s1 += a.x;
            s2 += a.y;
        }
        const double cnt = (double)win_count(x, g.X) * win_count(y, g.Y) * win_count(z, g.Z);
        const double mean = s1 / cnt;
        double var = s2 / cnt - mean * mean;
        if (var < 0.0) var = 0.0;
        out[v] = (float)(((double)f[v] - mean) / (sqrt(var) + 1e-6));
    }
}

// ---- conv1: 1 -> 16, ReLU (VALU) --------------------------------------------

__global__ void __launch_bounds__(kT)
conv1_kernel(const float* __restrict__ xin, Geo g, const float* __restrict__ w1 /*[27][16] bf16-rounded*/,
             const float* __restrict__ b1, uint16_t* __restrict__ a1) {
    __shared__ float sw[27 * 16 + 16];
    for (int i = threadIdx.x; i < 27 * 16; i += kT) sw[i] = w1[i];
    if (threadIdx.x < 16) sw[27 * 16 + threadIdx.x] = b1[threadIdx.x];
    __syncthreads();
    const int64_t n = g.verts();
    for (int64_t v = (int64_t)blockIdx.x * kT + threadIdx.x; v < n; v += (int64_t)gridDim.x * kT) {
        const int x = (int)(v % g.X), y = (int)((v / g.X) % g.Y), z = (int)(v / g.sxy);
        float nv[27];
#pragma unroll
        for (int t = 0; t < 27; t++) {
            const int xx = x + tap_dx(t), yy = y + tap_dy(t), zz = z + tap_dz(t);
            const bool ok = xx >= 0 && xx < g.X && yy >= 0 && yy < g.Y && zz >= 0 && zz < g.Z;
            nv[t] = ok ? xin[v + tap_dx(t) + (int64_t)tap_dy(t) * g.X + (int64_t)tap_dz(t) * g.sxy] : 0.f;
        }
        uint32_t packed[8];
#pragma unroll
        for (int c = 0; c < 16; c += 2) {
            float z0 = sw[27 * 16 + c], z1 = sw[27 * 16 + c + 1];
#pragma unroll
            for (int t = 0; t < 27; t++) {
                z0 = fmaf(sw[t * 16 + c], nv[t], z0);
                z1 = fmaf(sw[t * 16 + c + 1], nv[t], z1);
            }
            z0 = z0 > 0.f ? z0 : 0.f;
            z1 = z1 > 0.f ? z1 : 0.f;
            packed[c / 2] = (uint32_t)f2bf(z0) | ((uint32_t)f2bf(z1) << 16);
        }
        uint4* o = reinterpret_cast<uint4*>(a1 + v * 16);
        o[0] = make_uint4(packed[0], packed[1], packed[2], packed[3]);
        o[1] = make_uint4(packed[4], packed[5], packed[6], packed[7]);
    }
}

// ---- MFMA convolutions ------------------------------------------------------

enum ConvMode : int {
    kFwd = 0,      // out = relu(acc + bias), BF16
    kBwd = 1,      // out = acc * (gate > 0), BF16 (transposed convolution of a gradient)
    kFwdHead = 2,  // prob = sigmoid(head . relu(acc + bias)), FP32 (conv3 only)
};

__device__ __forceinline__ bf16x8 load8(const uint16_t* p) {
    const uint4 u = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(bf16x8, u);
}

// conv 16 -> 32: one wave computes 32 consecutive x-voxels x 32 output
// channels with one 32x32x16 MFMA per tap.  Operand layouts (wave64):
// A[row = lane%32][k = 8*(lane/32) .. +8], B[k = 8*(lane/32) .. +8][col = lane%32],
// D[i]: row 8*(i/4) + 4*(lane/32) + i%4, col lane%32.
template <int kMode>
__global__ void __launch_bounds__(kT)
conv16to32_kernel(const uint16_t* __restrict__ in, Geo g, const uint16_t* __restrict__ w /*[27][32][16]*/,
                  const float* __restrict__ bias, const uint16_t* __restrict__ gate, uint16_t* __restrict__ out) {
    __shared__ __attribute__((aligned(16))) uint16_t sw[27 * 32 * 16];
    __shared__ float sb[32];
    for (int i = threadIdx.x; i < 27 * 32 * 16 / 8; i += kT)
        reinterpret_cast<uint4*>(sw)[i] = reinterpret_cast<const uint4*>(w)[i];
    if (kMode == kFwd && threadIdx.x < 32) sb[threadIdx.x] = bias[threadIdx.x];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 31, h = lane >> 5;
    const int tiles_x = (g.X + 31) / 32;
    const int64_t ntiles = (int64_t)tiles_x * g.Y * g.Z;
    for (int64_t tile = (int64_t)blockIdx.x * 4 + wave; tile < ntiles; tile += (int64_t)gridDim.x * 4) {
        const int x0 = (int)(tile % tiles_x) * 32;
        const int y = (int)((tile / tiles_x) % g.Y), z = (int)(tile / ((int64_t)tiles_x * g.Y));
        f32x16 acc = {};
        const int x = x0 + r;
#pragma unroll 3
        for (int t = 0; t < 27; t++) {
            const int xx = x + tap_dx(t), yy = y + tap_dy(t), zz = z + tap_dz(t);
            const bool ok = x < g.X && xx >= 0 && xx < g.X && yy >= 0 && yy < g.Y && zz >= 0 && zz < g.Z;
            bf16x8 a = {};
            if (ok) a = load8(in + ((int64_t)xx + (int64_t)yy * g.X + (int64_t)zz * g.sxy) * 16 + 8 * h);
            const bf16x8 b = load8(sw + (t * 32 + r) * 16 + 8 * h);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        const int co = r;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int row = 8 * (i / 4) + 4 * h + (i % 4);
            const int xo = x0 + row;
            if (xo >= g.X) continue;
            const int64_t vo = (int64_t)xo + (int64_t)y * g.X + (int64_t)z * g.sxy;
            float val = acc[i];
            if (kMode == kFwd) {
                val += sb[co];
                val = val > 0.f ? val : 0.f;
            } else {
                if (!(bf2f(gate[vo * 32 + co]) > 0.f)) val = 0.f;
            }
            out[vo * 32 + co] = f2bf(val);
        }
    }
}

// conv 32 -> 16: one wave computes 64 consecutive x-voxels (four 16-voxel
// sub-tiles) x 16 output channels, one 16x16x32 MFMA per sub-tile and tap.
// A[row = lane%16][k = 8*(lane/16) .. +8], B[k = 8*(lane/16) .. +8][col = lane%16],
// D[i]: row 4*(lane/16) + i, col lane%16.
template <int kMode>
__global__ void __launch_bounds__(kT)
conv32to16_kernel(const uint16_t* __restrict__ in, Geo g, const uint16_t* __restrict__ w /*[27][16][32]*/,
                  const float* __restrict__ bias, const float* __restrict__ head /*[16] + bias*/,
                  const uint16_t* __restrict__ gate, uint16_t* __restrict__ out, float* __restrict__ prob) {
    __shared__ __attribute__((aligned(16))) uint16_t sw[27 * 16 * 32];
    __shared__ float sb[16], sh[17];
    for (int i = threadIdx.x; i < 27 * 16 * 32 / 8; i += kT)
        reinterpret_cast<uint4*>(sw)[i] = reinterpret_cast<const uint4*>(w)[i];
    if (kMode != kBwd && threadIdx.x < 16) sb[threadIdx.x] = bias[threadIdx.x];
    if (kMode == kFwdHead && threadIdx.x < 17) sh[threadIdx.x] = head[threadIdx.x];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, q = lane >> 4;
    const int tiles_x = (g.X + 63) / 64;
    const int64_t ntiles = (int64_t)tiles_x * g.Y * g.Z;
    for (int64_t tile = (int64_t)blockIdx.x * 4 + wave; tile < ntiles; tile += (int64_t)gridDim.x * 4) {
        const int x0 = (int)(tile % tiles_x) * 64;
        const int y = (int)((tile / tiles_x) % g.Y), z = (int)(tile / ((int64_t)tiles_x * g.Y));
        f32x4 acc[4] = {};
#pragma unroll 3
        for (int t = 0; t < 27; t++) {
            const bf16x8 b = load8(sw + (t * 16 + r) * 32 + 8 * q);
            const int yy = y + tap_dy(t), zz = z + tap_dz(t);
            const bool rowok = yy >= 0 && yy < g.Y && zz >= 0 && zz < g.Z;
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const int x = x0 + 16 * s + r, xx = x + tap_dx(t);
                bf16x8 a = {};
                if (rowok && x < g.X && xx >= 0 && xx < g.X)
                    a = load8(in + ((int64_t)xx + (int64_t)yy * g.X + (int64_t)zz * g.sxy) * 32 + 8 * q);
                acc[s] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[s], 0, 0, 0);
            }
        }
        const int co = r;
#pragma unroll
        for (int s = 0; s < 4; s++) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int xo = x0 + 16 * s + 4 * q + i;
                const bool in_range = xo < g.X;
                const int64_t vo = (int64_t)xo + (int64_t)y * g.X + (int64_t)z * g.sxy;
                float val = acc[s][i];
                if (kMode == kBwd) {
                    if (in_range && !(bf2f(gate[vo * 16 + co]) > 0.f)) val = 0.f;
                    if (in_range) out[vo * 16 + co] = f2bf(val);
                } else {
                    val += sb[co];
                    val = val > 0.f ? val : 0.f;
                    if (kMode == kFwd) {
                        if (in_range) out[vo * 16 + co] = f2bf(val);
                    } else {
                        // head: the 16 channels of a voxel sit in 16 consecutive lanes
                        float p = sh[co] * bf2f(f2bf(val));
                        p += __shfl_xor(p, 1);
                        p += __shfl_xor(p, 2);
                        p += __shfl_xor(p, 4);
                        p += __shfl_xor(p, 8);
                        if (in_range && co == 0) prob[vo] = 1.f / (1.f + expf(-(p + sh[16])));
                    }
                }
            }
        }
    }
}

// ---- head + loss (training) -------------------------------------------------

// Per voxel: logit = bh + sum_c wh[c] a3[c], p = sigmoid(logit); weighted BCE
// with weight pos_w on targets 1; dlogit = w (p - t) / n; dz3 = dlogit wh (a3 > 0).
// Per-block partials: [loss, dbh, dwh[16]].
constexpr int kHeadPart = 18;

__global__ void __launch_bounds__(kT)
head_loss_kernel(const uint16_t* __restrict__ a3, Geo g, const float* __restrict__ head, const uint8_t* __restrict__ target,
                 float pos_w, float inv_n, uint16_t* __restrict__ dz3, float* __restrict__ part) {
    __shared__ float red[kT / 64][kHeadPart];
    float acc[kHeadPart] = {};
    float wh[16];
#pragma unroll
    for (int c = 0; c < 16; c++) wh[c] = head[c];
    const float bh = head[16];
    const int64_t n = g.verts();
    for (int64_t v = (int64_t)blockIdx.x * kT + threadIdx.x; v < n; v += (int64_t)gridDim.x * kT) {
        float a[16];
        const uint4* p4 = reinterpret_cast<const uint4*>(a3 + v * 16);
        const uint4 u0 = p4[0], u1 = p4[1];
        const uint32_t w8[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
#pragma unroll
        for (int c = 0; c < 8; c++) {
            a[2 * c] = bf2f((uint16_t)(w8[c] & 0xFFFFu));
            a[2 * c + 1] = bf2f((uint16_t)(w8[c] >> 16));
        }
        float logit = bh;
#pragma unroll
        for (int c = 0; c < 16; c++) logit = fmaf(wh[c], a[c], logit);
        const float t = target[v] ? 1.f : 0.f;
        const float wgt = target[v] ? pos_w : 1.f;
        // log-sigmoid forms, stable for large |logit|
        const float lsp = logit >= 0.f ? -log1pf(expf(-logit)) : logit - log1pf(expf(logit));  // log p
        const float lsn = lsp - logit;                                                          // log (1-p)
        acc[0] += -wgt * (t * lsp + (1.f - t) * lsn);
        const float p = 1.f / (1.f + expf(-logit));
        const float dl = wgt * (p - t) * inv_n;
        acc[1] += dl;
        uint32_t o[8];
#pragma unroll
        for (int c = 0; c < 16; c += 2) {
            acc[2 + c] += dl * a[c];
            acc[3 + c] += dl * a[c + 1];
            const float d0 = a[c] > 0.f ? dl * wh[c] : 0.f, d1 = a[c + 1] > 0.f ? dl * wh[c + 1] : 0.f;
            o[c / 2] = (uint32_t)f2bf(d0) | ((uint32_t)f2bf(d1) << 16);
        }
        uint4* q4 = reinterpret_cast<uint4*>(dz3 + v * 16);
        q4[0] = make_uint4(o[0], o[1], o[2], o[3]);
        q4[1] = make_uint4(o[4], o[5], o[6], o[7]);
    }
    // block reduction in a fixed order
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kHeadPart; k++) {
        float x = acc[k];
        for (int m = 32; m >= 1; m >>= 1) x += __shfl_xor(x, m);
        if (lane == 0) red[wave][k] = x;
    }
    __syncthreads();
    if (threadIdx.x < kHeadPart) {
        float x = 0.f;
        for (int w = 0; w < kT / 64; w++) x += red[w][threadIdx.x];
        part[(int64_t)blockIdx.x * kHeadPart + threadIdx.x] = x;
    }
}

// ---- weight gradients (VALU) ------------------------------------------------

// dW[t][co][ci] = sum_v a_in(v + off_t, ci) dz(v, co), db[co] = sum_v dz(v, co).
// A block walks 64-voxel x-segments; the input halo (9 rows x 66 voxels) and
// the gradient segment are staged in LDS.  Thread (co, pair of ci) keeps 27x2
// accumulators; per-block partials [Cout*27*Cin + Cout] are reduced later.
template <int kCin, int kCout>
__global__ void __launch_bounds__(kT)
wgrad_kernel(const uint16_t* __restrict__ ain, const uint16_t* __restrict__ dz, Geo g, float* __restrict__ part) {
    static_assert(kCout * kCin / 2 == kT, "one thread per (co, ci pair)");
    constexpr int kSeg = 64, kW = kSeg + 2;
    __shared__ __attribute__((aligned(16))) uint16_t sa[9 * kW * kCin];
    __shared__ __attribute__((aligned(16))) uint16_t sd[kSeg * kCout];
    const int co = threadIdx.x / (kCin / 2), cp = threadIdx.x % (kCin / 2);
    float acc[27][2] = {};
    float bacc = 0.f;
    const int segs_x = (g.X + kSeg - 1) / kSeg;
    const int64_t nseg = (int64_t)segs_x * g.Y * g.Z;
    for (int64_t sg = blockIdx.x; sg < nseg; sg += gridDim.x) {
        const int x0 = (int)(sg % segs_x) * kSeg;
        const int y = (int)((sg / segs_x) % g.Y), z = (int)(sg / ((int64_t)segs_x * g.Y));
        __syncthreads();
        // stage the input halo: rows (dy, dz), voxels x0-1 .. x0+64, all channels (4-byte words)
        constexpr int kWords = 9 * kW * kCin / 2;
        for (int i = threadIdx.x; i < kWords; i += kT) {
            const int c2 = i % (kCin / 2), rest = i / (kCin / 2);
            const int xi = rest % kW, row = rest / kW;
            const int yy = y + row % 3 - 1, zz = z + row / 3 - 1, xx = x0 - 1 + xi;
            uint32_t val = 0u;
            if (xx >= 0 && xx < g.X && yy >= 0 && yy < g.Y && zz >= 0 && zz < g.Z)
                val = reinterpret_cast<const uint32_t*>(ain + ((int64_t)xx + (int64_t)yy * g.X + (int64_t)zz * g.sxy) * kCin)[c2];
            reinterpret_cast<uint32_t*>(sa)[i] = val;
        }
        for (int i = threadIdx.x; i < kSeg * kCout / 2; i += kT) {
            const int c2 = i % (kCout / 2), xi = i / (kCout / 2), xx = x0 + xi;
            uint32_t val = 0u;
            if (xx < g.X) val = reinterpret_cast<const uint32_t*>(dz + ((int64_t)xx + (int64_t)y * g.X + (int64_t)z * g.sxy) * kCout)[c2];
            reinterpret_cast<uint32_t*>(sd)[i] = val;
        }
        __syncthreads();
        for (int xi = 0; xi < kSeg; xi++) {
            const float d = bf2f(sd[xi * kCout + co]);
            if (d == 0.f) continue;
            bacc += d;
#pragma unroll
            for (int t = 0; t < 27; t++) {
                const int row = (tap_dy(t) + 1) + 3 * (tap_dz(t) + 1);
                const uint32_t pr = reinterpret_cast<const uint32_t*>(sa)[((row * kW) + xi + 1 + tap_dx(t)) * (kCin / 2) + cp];
                acc[t][0] = fmaf(d, bf2f((uint16_t)(pr & 0xFFFFu)), acc[t][0]);
                acc[t][1] = fmaf(d, bf2f((uint16_t)(pr >> 16)), acc[t][1]);
            }
        }
    }
    float* out = part + (int64_t)blockIdx.x * (27 * kCout * kCin + kCout);
#pragma unroll
    for (int t = 0; t < 27; t++) {
        out[(t * kCout + co) * kCin + 2 * cp] = acc[t][0];
        out[(t * kCout + co) * kCin + 2 * cp + 1] = acc[t][1];
    }
    if (cp == 0) out[27 * kCout * kCin + co] = bacc;
}

// conv1 gradient: dW1[t][co] = sum_v n(v + off_t) dz1(v, co), db1[co] = sum_v dz1(v, co).
__global__ void __launch_bounds__(kT)
wgrad1_kernel(const float* __restrict__ xin, const uint16_t* __restrict__ dz1, Geo g, float* __restrict__ part) {
    constexpr int kSeg = 64, kW = kSeg + 2;
    __shared__ float sa[9 * kW];
    __shared__ float sd[kSeg * 16];
    const int co = threadIdx.x % 16, tg = threadIdx.x / 16;  // taps tg and tg + 16
    float acc0 = 0.f, acc1 = 0.f, bacc = 0.f;
    const int segs_x = (g.X + kSeg - 1) / kSeg;
    const int64_t nseg = (int64_t)segs_x * g.Y * g.Z;
    for (int64_t sg = blockIdx.x; sg < nseg; sg += gridDim.x) {
        const int x0 = (int)(sg % segs_x) * kSeg;
        const int y = (int)((sg / segs_x) % g.Y), z = (int)(sg / ((int64_t)segs_x * g.Y));
        __syncthreads();
        for (int i = threadIdx.x; i < 9 * kW; i += kT) {
            const int xi = i % kW, row = i / kW;
            const int yy = y + row % 3 - 1, zz = z + row / 3 - 1, xx = x0 - 1 + xi;
            float val = 0.f;
            if (xx >= 0 && xx < g.X && yy >= 0 && yy < g.Y && zz >= 0 && zz < g.Z)
                val = xin[(int64_t)xx + (int64_t)yy * g.X + (int64_t)zz * g.sxy];
            sa[i] = val;
        }
        for (int i = threadIdx.x; i < kSeg * 16; i += kT) {
            const int c = i % 16, xi = i / 16, xx = x0 + xi;
            sd[i] = xx < g.X ? bf2f(dz1[((int64_t)xx + (int64_t)y * g.X + (int64_t)z * g.sxy) * 16 + c]) : 0.f;
        }
        __syncthreads();
        const int t0 = tg, t1 = tg + 16;
        for (int xi = 0; xi < kSeg; xi++) {
            const float d = sd[xi * 16 + co];
            if (tg == 0) bacc += d;
            {
                const int row = (tap_dy(t0) + 1) + 3 * (tap_dz(t0) + 1);
                acc0 = fmaf(d, sa[row * kW + xi + 1 + tap_dx(t0)], acc0);
            }
            if (t1 < 27) {
                const int row = (tap_dy(t1) + 1) + 3 * (tap_dz(t1) + 1);
                acc1 = fmaf(d, sa[row * kW + xi + 1 + tap_dx(t1)], acc1);
            }
        }
    }
    float* out = part + (int64_t)blockIdx.x * (27 * 16 + 16);
    out[tg * 16 + co] = acc0;
    if (tg + 16 < 27) out[(tg + 16) * 16 + co] = acc1;
    if (tg == 0) out[27 * 16 + co] = bacc;
}

// Sum of per-block partials (fixed order): out[i] = sum_b part[b][i].
__global__ void __launch_bounds__(kT)
reduce_parts_kernel(const float* __restrict__ part, int nblocks, int width, float* __restrict__ out) {
    for (int i = blockIdx.x * kT + threadIdx.x; i < width; i += gridDim.x * kT) {
        double s = 0.0;
        for (int b = 0; b < nblocks; b++) s += (double)part[(int64_t)b * width + i];
        out[i] = (float)s;
    }
}

// [loss, dbh, dwh[16]] -> gradient layout (wh[16], bh at kOffWh)
__global__ void reorder_head_kernel(const float* __restrict__ red, float* __restrict__ grad) {
    const int c = threadIdx.x;
    if (c < 16) grad[kOffWh + c] = red[2 + c];
    if (c == 16) grad[kOffWh + 16] = red[1];
}

// ---- parameters ---------------------------------------------------------------

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ gr, float* __restrict__ m,
                            float* __restrict__ v, int n, float lr, float b1t, float b2t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = gr[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.f - b1t), vh = vi / (1.f - b2t);
    p[i] -= lr * mh / (sqrtf(vh) + 1e-8f);
}

// Packed operands from the FP32 parameters.
__global__ void pack_kernel(const float* __restrict__ p, CnnPacked k) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kW1) k.w1[i] = bf2f(f2bf(p[kOffW1 + i]));
    if (i < 16) k.b1[i] = p[kOffB1 + i];
    if (i < kW2) {
        // W2 [t][co 32][ci 16] and its mirror-transpose W2T [t'][ci 16][co 32] (t' = 26 - t)
        const uint16_t h = f2bf(p[kOffW2 + i]);
        k.w2[i] = h;
        const int t = i / 512, co = (i / 16) % 32, ci = i % 16;
        k.w2t[((26 - t) * 16 + ci) * 32 + co] = h;
    }
    if (i < 32) k.b2[i] = p[kOffB2 + i];
    if (i < kW3) {
        // W3 [t][co 16][ci 32] and W3T [t'][ci 32][co 16]
        const uint16_t h = f2bf(p[kOffW3 + i]);
        k.w3[i] = h;
        const int t = i / 512, co = (i / 32) % 16, ci = i % 32;
        k.w3t[((26 - t) * 32 + ci) * 16 + co] = h;
    }
    if (i < 16) k.b3[i] = p[kOffB3 + i];
    if (i < 17) k.head[i] = p[kOffWh + i];  // wh[16], bh
}

// ---- mask from probabilities: threshold, then Chebyshev dilation ------------

__global__ void threshold_kernel(const float* __restrict__ prob, int64_t n, float thr, uint8_t* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = prob[i] > thr ? 1 : 0;
}

// One axis of a max filter of radius r.
__global__ void dilate_axis_kernel(const uint8_t* __restrict__ in, Geo g, int axis, int r, uint8_t* __restrict__ out) {
    const int64_t n = g.verts();
    const int64_t stride = axis == 0 ? 1 : (axis == 1 ? (int64_t)g.X : g.sxy);
    const int len = axis == 0 ? g.X : (axis == 1 ? g.Y : g.Z);
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const int c = axis == 0 ? (int)(v % g.X) : (axis == 1 ? (int)((v / g.X) % g.Y) : (int)(v / g.sxy));
        uint8_t m = 0;
        for (int d = -r; d <= r; d++) {
            const int cc = c + d;
            if (cc < 0 || cc >= len) continue;
            m |= in[v + d * stride];
        }
        out[v] = m;
    }
}

__global__ void count_kernel(const uint8_t* __restrict__ m, int64_t n, unsigned long long* __restrict__ out) {
    uint32_t c = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        c += m[i] ? 1u : 0u;
    for (int s = 32; s >= 1; s >>= 1) c += __shfl_xor(c, s);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, (unsigned long long)c);
}

inline uint32_t blocks_for(int64_t n, uint32_t cap = 16384) {
    const int64_t b = (n + kT - 1) / kT;
    return (uint32_t)(b < 1 ? 1 : (b > cap ? cap : b));
}

// SplitMix64 -> uniform (0,1) -> normal (Box-Muller), for the initial weights.
struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return ((double)(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal() {
        const double u1 = uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

}  // namespace

// ---- host side -----------------------------------------------------------------

std::vector<float> cnn_initial_params(uint64_t seed) {
    std::vector<float> p(kNumParams, 0.f);
    Rng rng{seed};
    auto he = [&](int off, int n, int fan_in) {
        const double s = std::sqrt(2.0 / fan_in);
        for (int i = 0; i < n; i++) p[off + i] = (float)(s * rng.normal());
    };
    he(kOffW1, kW1, 27);
    he(kOffW2, kW2, 27 * 16);
    he(kOffW3, kW3, 27 * 32);
    for (int i = 0; i < 16; i++) p[kOffWh + i] = (float)(std::sqrt(1.0 / 16) * rng.normal());
    p[kOffWh + 16] = -1.0f;  // start below 0.5: most voxels are negatives
    return p;
}

hipError_t CnnState::ensure_volume(int64_t verts) {
    if (verts <= vol_cap) return hipSuccess;
    hipError_t e;
    if ((e = tmp1.ensure(sizeof(double2) * verts)) != hipSuccess) return e;
    if ((e = tmp2.ensure(sizeof(double2) * verts)) != hipSuccess) return e;
    if ((e = xin.ensure(sizeof(float) * verts)) != hipSuccess) return e;
    if ((e = a1.ensure(2 * 16 * verts)) != hipSuccess) return e;
    if ((e = a2.ensure(2 * 32 * verts)) != hipSuccess) return e;
    if ((e = a3.ensure(2 * 16 * verts)) != hipSuccess) return e;
    if ((e = prob.ensure(sizeof(float) * verts)) != hipSuccess) return e;
    if ((e = m0.ensure(verts)) != hipSuccess) return e;
    if ((e = m1.ensure(verts)) != hipSuccess) return e;
    vol_cap = verts;
    return hipSuccess;
}

hipError_t CnnState::ensure_train(int64_t verts) {
    hipError_t e;
    if ((e = dz3.ensure(2 * 16 * verts)) != hipSuccess) return e;
    if ((e = dz2.ensure(2 * 32 * verts)) != hipSuccess) return e;
    if ((e = dz1.ensure(2 * 16 * verts)) != hipSuccess) return e;
    return hipSuccess;
}

void CnnState::release_all() {
    DevBuf* all[] = {&params, &grads, &m, &v, &pk, &parts, &loss, &tmp1, &tmp2, &xin, &a1, &a2, &a3,
                     &prob, &m0, &m1, &dz3, &dz2, &dz1};
    for (DevBuf* b : all) b->release();
    vol_cap = 0;
}

CnnPacked CnnState::packed() const {
    CnnPacked k;
    uint8_t* base = pk.as<uint8_t>();
    k.w1 = reinterpret_cast<float*>(base);
    k.b1 = k.w1 + kW1;
    k.b2 = k.b1 + 16;
    k.b3 = k.b2 + 32;
    k.head = k.b3 + 16;
    uint16_t* h = reinterpret_cast<uint16_t*>(base + 4096);
    k.w2 = h;
    k.w2t = h + kW2;
    k.w3 = h + 2 * kW2;
    k.w3t = h + 2 * kW2 + kW3;
    return k;
}

hipError_t cnn_pack(CnnState& s, hipStream_t st) {
    hipLaunchKernelGGL(pack_kernel, dim3((kW2 + kT - 1) / kT), dim3(kT), 0, st, s.params.as<float>(), s.packed());
    return hipGetLastError();
}

hipError_t cnn_forward(CnnState& s, const float* values, int X, int Y, int Z, bool keep_a3, hipStream_t st) {
    const Geo g{X, Y, Z, (int64_t)X * Y};
    const int64_t n = g.verts();
    const CnnPacked k = s.packed();
    hipLaunchKernelGGL(norm_x_kernel, dim3(blocks_for(n)), dim3(kT), 0, st, values, g, s.tmp1.as<double2>());
    hipLaunchKernelGGL(norm_y_kernel, dim3(blocks_for(n)), dim3(kT), 0, st, s.tmp1.as<double2>(), g, s.tmp2.as<double2>());
    hipLaunchKernelGGL(norm_z_kernel, dim3(blocks_for(n)), dim3(kT), 0, st, s.tmp2.as<double2>(), values, g,
                       s.xin.as<float>());
    hipLaunchKernelGGL(conv1_kernel, dim3(blocks_for(n)), dim3(kT), 0, st, s.xin.as<float>(), g, k.w1, k.b1,
                       s.a1.as<uint16_t>());
    const int64_t t32 = (int64_t)((X + 31) / 32) * Y * Z, t64 = (int64_t)((X + 63) / 64) * Y * Z;
    const uint32_t b32 = (uint32_t)std::min<int64_t>((t32 + 3) / 4, 65536), b64 = (uint32_t)std::min<int64_t>((t64 + 3) / 4, 65536);
    hipLaunchKernelGGL(conv16to32_kernel<kFwd>, dim3(b32), dim3(kT), 0, st, s.a1.as<uint16_t>(), g, k.w2, k.b2,
                       (const uint16_t*)nullptr, s.a2.as<uint16_t>());
    if (keep_a3)
        hipLaunchKernelGGL(conv32to16_kernel<kFwd>, dim3(b64), dim3(kT), 0, st, s.a2.as<uint16_t>(), g, k.w3, k.b3,
                           k.head, (const uint16_t*)nullptr, s.a3.as<uint16_t>(), (float*)nullptr);
    else
        hipLaunchKernelGGL(conv32to16_kernel<kFwdHead>, dim3(b64), dim3(kT), 0, st, s.a2.as<uint16_t>(), g, k.w3,
                           k.b3, k.head, (const uint16_t*)nullptr, (uint16_t*)nullptr, s.prob.as<float>());
    return hipGetLastError();
}

hipError_t cnn_backward(CnnState& s, int X, int Y, int Z, const uint8_t* target, float pos_w, hipStream_t st) {
    const Geo g{X, Y, Z, (int64_t)X * Y};
    const int64_t n = g.verts();
    const CnnPacked k = s.packed();
    const uint32_t hb = blocks_for(n, 1024);
    hipLaunchKernelGGL(head_loss_kernel, dim3(hb), dim3(kT), 0, st, s.a3.as<uint16_t>(), g, k.head, target, pos_w,
                       (float)(1.0 / (double)n), s.dz3.as<uint16_t>(), s.parts.as<float>());
    float* grad = s.grads.as<float>();
    float* loss = s.loss.as<float>();
    // head: [loss, dbh, dwh[16]] -> loss, grads at kOffWh
    hipLaunchKernelGGL(reduce_parts_kernel, dim3(1), dim3(kT), 0, st, s.parts.as<float>(), (int)hb, kHeadPart,
                       loss);
    hipLaunchKernelGGL(reorder_head_kernel, dim3(1), dim3(32), 0, st, loss, grad);
    const int64_t t32 = (int64_t)((X + 31) / 32) * Y * Z, t64 = (int64_t)((X + 63) / 64) * Y * Z;
    const uint32_t b32 = (uint32_t)std::min<int64_t>((t32 + 3) / 4, 65536), b64 = (uint32_t)std::min<int64_t>((t64 + 3) / 4, 65536);
    // dz2 = (dz3 conv^T W3) * (a2 > 0): a 16 -> 32 convolution with W3T
    hipLaunchKernelGGL(conv16to32_kernel<kBwd>, dim3(b32), dim3(kT), 0, st, s.dz3.as<uint16_t>(), g, k.w3t,
                       (const float*)nullptr, s.a2.as<uint16_t>(), s.dz2.as<uint16_t>());
    // dz1 = (dz2 conv^T W2) * (a1 > 0): a 32 -> 16 convolution with W2T
    hipLaunchKernelGGL(conv32to16_kernel<kBwd>, dim3(b64), dim3(kT), 0, st, s.dz2.as<uint16_t>(), g, k.w2t,
                       (const float*)nullptr, (const float*)nullptr, s.a1.as<uint16_t>(), s.dz1.as<uint16_t>(),
                       (float*)nullptr);
    const int64_t nseg = (int64_t)((X + 63) / 64) * Y * Z;
    const int wb = (int)std::min<int64_t>(nseg, kGradBlocks);
    hipLaunchKernelGGL((wgrad_kernel<32, 16>), dim3(wb), dim3(kT), 0, st, s.a2.as<uint16_t>(), s.dz3.as<uint16_t>(), g,
                       s.parts.as<float>());
    hipLaunchKernelGGL(reduce_parts_kernel, dim3((kW3 + 16 + kT - 1) / kT), dim3(kT), 0, st, s.parts.as<float>(), wb,
                       kW3 + 16, grad + kOffW3);
    hipLaunchKernelGGL((wgrad_kernel<16, 32>), dim3(wb), dim3(kT), 0, st, s.a1.as<uint16_t>(), s.dz2.as<uint16_t>(), g,
                       s.parts.as<float>());
    hipLaunchKernelGGL(reduce_parts_kernel, dim3((kW2 + 32 + kT - 1) / kT), dim3(kT), 0, st, s.parts.as<float>(), wb,
                       kW2 + 32, grad + kOffW2);
    hipLaunchKernelGGL(wgrad1_kernel, dim3(wb), dim3(kT), 0, st, s.xin.as<float>(), s.dz1.as<uint16_t>(), g,
                       s.parts.as<float>());
    hipLaunchKernelGGL(reduce_parts_kernel, dim3((kW1 + 16 + kT - 1) / kT), dim3(kT), 0, st, s.parts.as<float>(), wb,
                       kW1 + 16, grad + kOffW1);
    return hipGetLastError();
}

hipError_t cnn_adam(CnnState& s, float lr, hipStream_t st) {
    s.step++;
    const float b1t = std::pow(0.9f, (float)s.step), b2t = std::pow(0.999f, (float)s.step);
    hipLaunchKernelGGL(adam_kernel, dim3((kNumParams + kT - 1) / kT), dim3(kT), 0, st, s.params.as<float>(),
                       s.grads.as<float>(), s.m.as<float>(), s.v.as<float>(), kNumParams, lr, b1t, b2t);
    return cnn_pack(s, st);
}

hipError_t cnn_mask(CnnState& s, int X, int Y, int Z, float thr, int dilation, uint8_t* out,
                    unsigned long long* count, hipStream_t st) {
    const Geo g{X, Y, Z, (int64_t)X * Y};
    const int64_t n = g.verts();
    uint8_t* a = dilation > 0 ? s.m0.as<uint8_t>() : out;
    hipLaunchKernelGGL(threshold_kernel, dim3(blocks_for(n)), dim3(kT), 0, st, s.prob.as<float>(), n, thr, a);
    if (dilation > 0) {
        hipLaunchKernelGGL(dilate_axis_kernel, dim3(blocks_for(n)), dim3(kT), 0, st, s.m0.as<uint8_t>(), g, 0, dilation,
                           s.m1.as<uint8_t>());
        hipLaunchKernelGGL(dilate_axis_kernel, dim3(blocks_for(n)), dim3(kT), 0, st, s.m1.as<uint8_t>(), g, 1, dilation,
                           s.m0.as<uint8_t>());
        hipLaunchKernelGGL(dilate_axis_kernel, dim3(blocks_for(n)), dim3(kT), 0, st, s.m0.as<uint8_t>(), g, 2, dilation,
                           out);
    }
    if (count) hipLaunchKernelGGL(count_kernel, dim3(blocks_for(n, 4096)), dim3(kT), 0, st, out, n, count);
    return hipGetLastError();
}

}  // namespace msceer
