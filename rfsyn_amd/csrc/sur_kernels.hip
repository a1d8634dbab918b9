// Kernels of the learned surrogate (fdtd_surrogate.h): a BF16 matrix
// product on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16, FP32
// accumulation) with split-K, 3D convolution as im2col / col2im, GELU,
// and the Adam step.  No reference counterpart: the reference has no
// learned part.
//
// Activations are channel-last (NDHWC), so the im2col row of an output
// voxel is 27 contiguous runs of Cin values and the product's output
// [voxels, Cout] is the next layer's input layout.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "rf_device.h"
#include "rf_hip.h"

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even; NaN stays NaN.
__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return static_cast<uint16_t>((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}

// C(m, n) = sum_k A(m, k) B(k, n) with A(m, k) = A[m sam + k sak] and
// B(k, n) = B[k sbk + n sbn] (any transposition by strides), on the gfx950
// v_mfma_f32_16x16x32_bf16 instruction (BF16 operands, FP32 accumulation).
// Fragment layout (lane l): A row l%16, k 8(l/16)..+7; B column l%16,
// k 8(l/16)..+7; D column l%16, rows 4(l/16)..+3.  A block of 4
// wavefronts stacked along M computes a (64 WM) x (16 WN) tile, each
// wavefront WM x WN fragments; k advances 32 at a time.  The block
// stages A as rows of k and B transposed (columns of k) in LDS, so every
// fragment is one 16-byte LDS read; an operand whose k (or, for A, m; for
// B, n) stride is 1 is read 8 values at a time.  gridDim.z splits K:
// split z covers k in [z kchunk, (z+1) kchunk) and writes its partial
// tile to part + z M N (or, with one split, C + bias directly).  M tiles
// run along x, N tiles along y.
constexpr int TK = 32, KP = TK + 8;   // k per step, padded LDS row (80 bytes)

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef unsigned short v8u __attribute__((ext_vector_type(8)));

template <int WM, int WN>
__global__ void __launch_bounds__(256)
k_gemm(int M, int N, int K, const uint16_t *__restrict__ A, long long sam, long long sak,
       const uint16_t *__restrict__ B, long long sbk, long long sbn, float *__restrict__ C, long long ldc,
       const float *__restrict__ bias, int kchunk, float *__restrict__ part) {
    constexpr int BM = 64 * WM, BN = 16 * WN;
    __shared__ __align__(16) uint16_t sa[BM][KP];
    __shared__ __align__(16) uint16_t sb[BN][KP];   // B transposed: sb[n][k]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    const int kb = blockIdx.z * kchunk;
    const int ke = min(K, kb + kchunk);
    v4f acc[WM][WN];
#pragma unroll
    for (int i = 0; i < WM; i++)
#pragma unroll
        for (int j = 0; j < WN; j++) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
    // 8-wide global reads where the contiguous axis allows them
    const bool a_k8 = sak == 1 && (sam % 8) == 0 && (reinterpret_cast<uintptr_t>(A) % 16) == 0;
    const bool a_m8 = sam == 1 && (sak % 8) == 0 && (reinterpret_cast<uintptr_t>(A) % 16) == 0;
    const bool b_k8 = sbk == 1 && (sbn % 8) == 0 && (reinterpret_cast<uintptr_t>(B) % 16) == 0;
    const bool b_n8 = sbn == 1 && (sbk % 8) == 0 && (reinterpret_cast<uintptr_t>(B) % 16) == 0;
    for (int k0 = kb; k0 < ke; k0 += TK) {
        // ---- A tile [BM][TK]
        if (a_k8) {
            for (int e = tid; e < BM * (TK / 8); e += 256) {
                const int r = e / (TK / 8), c = (e % (TK / 8)) * 8;
                const int m = m0 + r, k = k0 + c;
                v8u v = v8u{0, 0, 0, 0, 0, 0, 0, 0};
                if (m < M && k + 7 < ke) {
                    v = *reinterpret_cast<const v8u *>(A + m * sam + k);
                } else if (m < M) {
#pragma unroll
                    for (int t = 0; t < 8; t++) v[t] = k + t < ke ? A[m * sam + (k + t)] : uint16_t(0);
                }
                *reinterpret_cast<v8u *>(&sa[r][c]) = v;
            }
        } else if (a_m8) {
            for (int e = tid; e < (BM / 8) * TK; e += 256) {
                const int r = (e % (BM / 8)) * 8, c = e / (BM / 8);
                const int m = m0 + r, k = k0 + c;
                v8u v = v8u{0, 0, 0, 0, 0, 0, 0, 0};
                if (k < ke && m + 7 < M) {
                    v = *reinterpret_cast<const v8u *>(A + m + k * sak);
                } else if (k < ke) {
#pragma unroll
                    for (int t = 0; t < 8; t++) v[t] = m + t < M ? A[(m + t) + k * sak] : uint16_t(0);
                }
#pragma unroll
                for (int t = 0; t < 8; t++) sa[r + t][c] = v[t];
            }
        } else {
            for (int e = tid; e < BM * TK; e += 256) {
                const int r = e / TK, c = e % TK;
                const int m = m0 + r, k = k0 + c;
                sa[r][c] = (m < M && k < ke) ? A[m * sam + k * sak] : uint16_t(0);
            }
        }
        // ---- B tile, transposed [BN][TK]
        if (b_k8) {
            for (int e = tid; e < BN * (TK / 8); e += 256) {
                const int r = e / (TK / 8), c = (e % (TK / 8)) * 8;
                const int n = n0 + r, k = k0 + c;
                v8u v = v8u{0, 0, 0, 0, 0, 0, 0, 0};
                if (n < N && k + 7 < ke) {
                    v = *reinterpret_cast<const v8u *>(B + k + n * sbn);
                } else if (n < N) {
#pragma unroll
                    for (int t = 0; t < 8; t++) v[t] = k + t < ke ? B[(k + t) + n * sbn] : uint16_t(0);
                }
                *reinterpret_cast<v8u *>(&sb[r][c]) = v;
            }
        } else if (b_n8 && BN >= 8) {
            for (int e = tid; e < (BN / 8) * TK; e += 256) {
                const int r = (e % (BN / 8)) * 8, c = e / (BN / 8);
                const int n = n0 + r, k = k0 + c;
                v8u v = v8u{0, 0, 0, 0, 0, 0, 0, 0};
                if (k < ke && n + 7 < N) {
                    v = *reinterpret_cast<const v8u *>(B + k * sbk + n);
                } else if (k < ke) {
#pragma unroll
                    for (int t = 0; t < 8; t++) v[t] = n + t < N ? B[k * sbk + (n + t)] : uint16_t(0);
                }
#pragma unroll
                for (int t = 0; t < 8; t++) sb[r + t][c] = v[t];
            }
        } else {
            for (int e = tid; e < BN * TK; e += 256) {
                const int r = e / TK, c = e % TK;
                const int n = n0 + r, k = k0 + c;
                sb[r][c] = (k < ke && n < N) ? B[k * sbk + n * sbn] : uint16_t(0);
            }
        }
        __syncthreads();
        const int kk = (lane >> 4) * 8;
        v8bf b[WN];
#pragma unroll
        for (int j = 0; j < WN; j++)
            b[j] = *reinterpret_cast<const v8bf *>(&sb[j * 16 + (lane & 15)][kk]);
#pragma unroll
        for (int i = 0; i < WM; i++) {
            const v8bf a = *reinterpret_cast<const v8bf *>(&sa[wave * 16 * WM + i * 16 + (lane & 15)][kk]);
#pragma unroll
            for (int j = 0; j < WN; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    float *out = gridDim.z > 1 ? part + (long long)blockIdx.z * M * N : C;
    const long long ld = gridDim.z > 1 ? N : ldc;
#pragma unroll
    for (int i = 0; i < WM; i++)
#pragma unroll
        for (int j = 0; j < WN; j++)
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int m = m0 + wave * 16 * WM + i * 16 + (lane >> 4) * 4 + t;
                const int n = n0 + j * 16 + (lane & 15);
                if (m < M && n < N) {
                    float v = acc[i][j][t];
                    if (gridDim.z == 1 && bias) v += bias[n];
                    out[m * ld + n] = v;
                }
            }
}

// Sum of the split partials in split order (deterministic), plus bias.
__global__ void k_split_sum(int M, int N, int splits, const float *__restrict__ part, float *__restrict__ C,
                            long long ldc, const float *__restrict__ bias) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)M * N) return;
    const int m = static_cast<int>(t / N), n = static_cast<int>(t % N);
    float s = 0.f;
    for (int z = 0; z < splits; z++) s += part[(long long)z * M * N + t];
    C[m * ldc + n] = s + (bias ? bias[n] : 0.f);
}

__global__ void k_to_bf16(const float *__restrict__ x, uint16_t *__restrict__ y, long long n) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        y[t] = to_bf16(x[t]);
}

// GELU with the erf form: a = 0.5 z (1 + erf(z / sqrt 2)); the activation
// is also written in BF16 as the next product's operand.
__global__ void k_gelu(const float *__restrict__ z, float *__restrict__ a, uint16_t *__restrict__ a16, long long n) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const float v = z[t];
        const float g = 0.5f * v * (1.f + erff(v * 0.70710678118654752f));
        if (a) a[t] = g;
        a16[t] = to_bf16(g);
    }
}

// dz = da * (Phi(z) + z phi(z)), Phi the normal CDF and phi its density.
__global__ void k_gelu_bwd(const float *__restrict__ z, const float *__restrict__ da, float *__restrict__ dz,
                           uint16_t *__restrict__ dz16, long long n) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const float v = z[t];
        const float cdf = 0.5f * (1.f + erff(v * 0.70710678118654752f));
        const float pdf = 0.39894228040143268f * expf(-0.5f * v * v);
        const float g = da[t] * (cdf + v * pdf);
        dz[t] = g;
        dz16[t] = to_bf16(g);
    }
}

// im2col of a 3x3x3 convolution with stride 2 and padding 1 over NDHWC
// input [B, S, S, S, C]: row (b, oz, oy, ox) of [B (S/2)^3, 27 C] holds,
// for kernel offset q = (kz 3 + ky) 3 + kx, the C input values at
// (2 oz + kz - 1, 2 oy + ky - 1, 2 ox + kx - 1), zero outside.
__global__ void k_im2col(const uint16_t *__restrict__ in, uint16_t *__restrict__ col, int B, int S, int C) {
    const int So = S / 2;
    const long long rows = (long long)B * So * So * So, K = 27LL * C;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < rows * K;
         t += (long long)gridDim.x * blockDim.x) {
        const long long r = t / K;
        const int kq = static_cast<int>(t % K);
        const int q = kq / C, c = kq % C;
        const int kz = q / 9, ky = (q / 3) % 3, kx = q % 3;
        const int ox = static_cast<int>(r % So), oy = static_cast<int>((r / So) % So),
                  oz = static_cast<int>((r / So / So) % So);
        const long long b = r / ((long long)So * So * So);
        const int z = 2 * oz + kz - 1, y = 2 * oy + ky - 1, x = 2 * ox + kx - 1;
        uint16_t v = 0;
        if (z >= 0 && z < S && y >= 0 && y < S && x >= 0 && x < S)
            v = in[(((b * S + z) * S + y) * S + x) * C + c];
        col[t] = v;
    }
}

// col2im as a gather (no atomics, fixed order): the input gradient of
// voxel (b, z, y, x, c) sums the column gradients of the output voxels
// whose window covers it: along each axis o = (z + 1 - k) / 2 for the
// kernel offsets k with z + 1 - k even and 0 <= o < S/2.
__global__ void k_col2im(const float *__restrict__ dcol, float *__restrict__ din, int B, int S, int C) {
    const int So = S / 2;
    const long long n = (long long)B * S * S * S * C;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const int c = static_cast<int>(t % C);
        long long r = t / C;
        const int x = static_cast<int>(r % S), y = static_cast<int>((r / S) % S), z = static_cast<int>((r / S / S) % S);
        const long long b = r / ((long long)S * S * S);
        float s = 0.f;
        for (int kz = 0; kz < 3; kz++) {
            const int pz = z + 1 - kz;
            if (pz < 0 || (pz & 1) || pz / 2 >= So) continue;
            for (int ky = 0; ky < 3; ky++) {
                const int py = y + 1 - ky;
                if (py < 0 || (py & 1) || py / 2 >= So) continue;
                for (int kx = 0; kx < 3; kx++) {
                    const int px = x + 1 - kx;
                    if (px < 0 || (px & 1) || px / 2 >= So) continue;
                    const long long row = ((b * So + pz / 2) * So + py / 2) * So + px / 2;
                    s += dcol[row * 27LL * C + ((kz * 3 + ky) * 3 + kx) * C + c];
                }
            }
        }
        din[t] = s;
    }
}

// Input voxels [B, S, S, S, C] (x, y, z, channel) from per-scene (x, z)
// maps [B, S, S, C] repeated along y (the scenes are invariant along y).
__global__ void k_expand_y(const float *__restrict__ maps, uint16_t *__restrict__ vox, int B, int S, int C) {
    const long long n = (long long)B * S * S * S * C;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const int c = static_cast<int>(t % C);
        long long r = t / C;
        const int z = static_cast<int>(r % S);
        r /= S;
        r /= S;   // y
        const int x = static_cast<int>(r % S);
        const long long b = r / S;
        vox[t] = to_bf16(maps[((b * S + x) * S + z) * C + c]);
    }
}

// h[bp, n] = U[bp / P, n] + V[bp, n] + bias[n] (first MLP layer: the
// scene part U computed once per scene, the query part V per probe).
__global__ void k_add_scene(const float *__restrict__ U, const float *__restrict__ V, const float *__restrict__ bias,
                            float *__restrict__ h, long long rows, int P, int N) {
    const long long n = rows * N;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const long long bp = t / N;
        const int c = static_cast<int>(t % N);
        h[t] = U[(bp / P) * N + c] + V[t] + bias[c];
    }
}

// dU[b, n] = sum over the scene's P queries of dh[b P + p, n].
__global__ void k_sum_queries(const float *__restrict__ dh, float *__restrict__ dU, uint16_t *__restrict__ dU16,
                              long long scenes, int P, int N) {
    const long long n = scenes * N;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const long long b = t / N;
        const int c = static_cast<int>(t % N);
        float s = 0.f;
        for (int p = 0; p < P; p++) s += dh[(b * P + p) * N + c];
        dU[t] = s;
        dU16[t] = to_bf16(s);
    }
}

// Mean squared error: d = (y - label) 2 / n into grad (and BF16), per
// block partial sums of (y - label)^2.
__global__ void k_mse(const float *__restrict__ y, const float *__restrict__ label, float *__restrict__ grad,
                      uint16_t *__restrict__ grad16, long long n, double *__restrict__ partials) {
    __shared__ double s[256];
    double acc = 0.0;
    const float scale = 2.f / static_cast<float>(n);
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const float d = y[t] - label[t];
        acc += static_cast<double>(d) * d;
        const float g = d * scale;
        grad[t] = g;
        grad16[t] = to_bf16(g);
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = s[0];
}

// Adam with bias correction (c1 = 1 - beta1^t, c2 = 1 - beta2^t); the
// FP32 master weights also refresh their BF16 copy.
__global__ void k_adam(float *__restrict__ w, uint16_t *__restrict__ w16, const float *__restrict__ g,
                       float *__restrict__ m, float *__restrict__ v, long long n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const float gt = g[t];
        const float mt = b1 * m[t] + (1.f - b1) * gt;
        const float vt = b2 * v[t] + (1.f - b2) * gt * gt;
        m[t] = mt;
        v[t] = vt;
        const float wt = w[t] - lr * (mt / c1) / (sqrtf(vt / c2) + eps);
        w[t] = wt;
        if (w16) w16[t] = to_bf16(wt);
    }
}

unsigned grid_for(long long n) {
    long long b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return static_cast<unsigned>(b);
}

int launched() { return hipGetLastError() == hipSuccess ? 0 : -1; }

}  // namespace

extern "C" {

int rf_sur_gemm(int M, int N, int K, const uint16_t *A, long long sam, long long sak, const uint16_t *B,
                long long sbk, long long sbn, float *C, long long ldc, const float *bias, int splits, float *part,
                rf_stream_t *s) {
    if (M <= 0 || N <= 0 || K <= 0 || splits < 1 || (splits > 1 && !part)) return -1;
    hipStream_t st = rf_raw_stream(s);
    int kchunk = (K + splits - 1) / splits;
    kchunk = (kchunk + TK - 1) / TK * TK;
    splits = (K + kchunk - 1) / kchunk;
    // tile shape by the width of the product: 256 x 16, 128 x 32 or 128 x 64
    const int bm = N <= 16 ? 256 : 128, bn = N <= 16 ? 16 : (N <= 32 ? 32 : 64);
    if ((N + bn - 1) / bn > 65535 || splits > 65535) return -1;
    const dim3 grd((M + bm - 1) / bm, (N + bn - 1) / bn, splits);
    if (N <= 16)
        hipLaunchKernelGGL((k_gemm<4, 1>), grd, dim3(256), 0, st, M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, bias,
                           kchunk, part);
    else if (N <= 32)
        hipLaunchKernelGGL((k_gemm<2, 2>), grd, dim3(256), 0, st, M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, bias,
                           kchunk, part);
    else
        hipLaunchKernelGGL((k_gemm<2, 4>), grd, dim3(256), 0, st, M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, bias,
                           kchunk, part);
    if (splits > 1)
        hipLaunchKernelGGL(k_split_sum, dim3(grid_for((long long)M * N)), dim3(256), 0, st, M, N, splits, part, C, ldc,
                           bias);
    return launched();
}

int rf_sur_to_bf16(const float *x, uint16_t *y, long long n, rf_stream_t *s) {
    hipLaunchKernelGGL(k_to_bf16, dim3(grid_for(n)), dim3(256), 0, rf_raw_stream(s), x, y, n);
    return launched();
}

int rf_sur_gelu(const float *z, float *a, uint16_t *a16, long long n, rf_stream_t *s) {
    hipLaunchKernelGGL(k_gelu, dim3(grid_for(n)), dim3(256), 0, rf_raw_stream(s), z, a, a16, n);
    return launched();
}

int rf_sur_gelu_bwd(const float *z, const float *da, float *dz, uint16_t *dz16, long long n, rf_stream_t *s) {
    hipLaunchKernelGGL(k_gelu_bwd, dim3(grid_for(n)), dim3(256), 0, rf_raw_stream(s), z, da, dz, dz16, n);
    return launched();
}

int rf_sur_im2col(const uint16_t *in, uint16_t *col, int B, int S, int C, rf_stream_t *s) {
    if (S % 2) return -1;
    const long long n = (long long)B * (S / 2) * (S / 2) * (S / 2) * 27 * C;
    hipLaunchKernelGGL(k_im2col, dim3(grid_for(n)), dim3(256), 0, rf_raw_stream(s), in, col, B, S, C);
    return launched();
}

int rf_sur_col2im(const float *dcol, float *din, int B, int S, int C, rf_stream_t *s) {
    if (S % 2) return -1;
    const long long n = (long long)B * S * S * S * C;
    hipLaunchKernelGGL(k_col2im, dim3(grid_for(n)), dim3(256), 0, rf_raw_stream(s), dcol, din, B, S, C);
    return launched();
}

int rf_sur_expand_y(const float *maps, uint16_t *vox, int B, int S, int C, rf_stream_t *s) {
    const long long n = (long long)B * S * S * S * C;
    hipLaunchKernelGGL(k_expand_y, dim3(grid_for(n)), dim3(256), 0, rf_raw_stream(s), maps, vox, B, S, C);
    return launched();
}

int rf_sur_add_scene(const float *U, const float *V, const float *bias, float *h, long long rows, int P, int N,
                     rf_stream_t *s) {
    hipLaunchKernelGGL(k_add_scene, dim3(grid_for(rows * N)), dim3(256), 0, rf_raw_stream(s), U, V, bias, h, rows, P,
                       N);
    return launched();
}

int rf_sur_sum_queries(const float *dh, float *dU, uint16_t *dU16, long long scenes, int P, int N, rf_stream_t *s) {
    hipLaunchKernelGGL(k_sum_queries, dim3(grid_for(scenes * N)), dim3(256), 0, rf_raw_stream(s), dh, dU, dU16, scenes,
                       P, N);
    return launched();
}

int rf_sur_mse(const float *y, const float *label, float *grad, uint16_t *grad16, long long n, double *partials,
               int n_partials, rf_stream_t *s) {
    hipLaunchKernelGGL(k_mse, dim3(n_partials), dim3(256), 0, rf_raw_stream(s), y, label, grad, grad16, n, partials);
    return launched();
}

int rf_sur_adam(float *w, uint16_t *w16, const float *g, float *m, float *v, long long n, float lr, float b1,
                float b2, float eps, float c1, float c2, rf_stream_t *s) {
    hipLaunchKernelGGL(k_adam, dim3(grid_for(n)), dim3(256), 0, rf_raw_stream(s), w, w16, g, m, v, n, lr, b1, b2, eps,
                       c1, c2);
    return launched();
}

}  // extern "C"
