// Dataset output stage on the device: a time-domain probe series as
// complex float samples (cli/fdtd_stage.c:20-24: re = (float)v, im = 0)
// and the reference's AWGN stage on them (engine/noise/awgn.c:16-44):
// the mean sample power summed in double in sample order, then per sample
// two PCG32 uniforms (engine/core/rng.c:3-23) and a float Box-Muller pair.
// The power sum and the generator follow the reference's order, so the
// power and the uniforms are its bits; logf, cosf and sinf are the
// device's, within a few float ulps of glibc's.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>

#include "rf_device.h"
#include "rf_hip.h"

namespace {

// Mean power of the samples, summed in sample order by one thread (the
// reference's order, so the double sum is its bits; a dataset example
// holds thousands of samples, not millions).
__global__ void k_iq_power(const double *__restrict__ td, long long n, double *__restrict__ power) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double p = 0.0;
    for (long long i = 0; i < n; i++) {
        const float mag = fabsf(static_cast<float>(td[i]) + 0.0f);   // cabsf(x + 0i) = |x|
        p += static_cast<double>(mag) * static_cast<double>(mag);
    }
    power[0] = n > 0 ? p / static_cast<double>(n) : 0.0;
}

// PCG32 state after k steps of its LCG from `state`: the constants of
// x -> a^k x + c (a^k - 1)/(a - 1) by binary powering (F. Brown, "Random
// number generation with arbitrary strides", Trans. Am. Nucl. Soc. 71,
// 1994).
__device__ __forceinline__ uint64_t lcg_advance(uint64_t state, uint64_t k, uint64_t inc) {
    uint64_t am = 1, ap = 0, cm = 6364136223846793005ULL, cp = inc;
    while (k) {
        if (k & 1u) {
            am *= cm;
            ap = ap * cm + cp;
        }
        cp = (cm + 1) * cp;
        cm *= cm;
        k >>= 1;
    }
    return am * state + ap;
}

// One PCG32 output from the state before the step (rng.c:14-20).
__device__ __forceinline__ uint32_t pcg_out(uint64_t old) {
    const uint32_t x = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = static_cast<uint32_t>(old >> 59u);
    return (x >> rot) | (x << ((0u - rot) & 31u));
}

__device__ __forceinline__ float pcg_uniform(uint64_t old) {
    return static_cast<float>(pcg_out(old) >> 8) / static_cast<float>(1u << 24);
}

// Sample i: I/Q from the series, plus (when noise) the pair from draws 2i
// and 2i + 1 of the stream whose state is `state0` (the seeded state,
// before its first draw).
__global__ void k_iq_noise(const double *__restrict__ td, long long n, int noise, float std_f, uint64_t state0,
                           uint64_t inc, float *__restrict__ out) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float re = static_cast<float>(td[i]) + 0.0f;
    float im = 0.0f;
    if (noise) {
        const uint64_t s1 = lcg_advance(state0, 2 * static_cast<uint64_t>(i), inc);
        const uint64_t s2 = s1 * 6364136223846793005ULL + inc;
        float u1 = pcg_uniform(s1);
        const float u2 = pcg_uniform(s2);
        if (u1 <= 0.0f) u1 = 1e-9f;
        const float r = sqrtf(-2.0f * logf(u1));
        const float w = 2.0f * 3.14159265358979323846f * u2;
        const float nre = std_f * r * cosf(w);
        const float nim = std_f * r * sinf(w);
        // samples[i] += nre + nim * I, in C's complex arithmetic
        re = re + (nre + nim * 0.0f);
        im = im + nim * 1.0f;
    }
    out[2 * i] = re;
    out[2 * i + 1] = im;
}

}  // namespace

// Device memory of the output stage's variant that allocates nothing
// (rf_iq_scratch): the mean power in the first 256 bytes, then up to
// RF_IQ_SCRATCH_SAMPLES I/Q samples.  Each GPU holds its own copy until the
// process ends; every call on a GPU shares it.
static __device__ __align__(256) unsigned char scratch[256 + 2 * RF_IQ_SCRATCH_SAMPLES * sizeof(float)];

extern "C" int rf_iq_scratch(double **power, float **iq) {
    // The array's name is a host-side stand-in; its address on the current
    // device comes from the runtime, once per device.
    static std::mutex mu;
    static void *addr[RF_MAX_DEVICES];
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= RF_MAX_DEVICES) return -1;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -1;
        addr[dev] = p;
    }
    *power = static_cast<double *>(addr[dev]);
    *iq = reinterpret_cast<float *>(static_cast<unsigned char *>(addr[dev]) + 256);
    return 0;
}

extern "C" int rf_launch_iq_power(const double *td, long long n, double *power, rf_stream_t *s) {
    if (n < 0 || (n && !td) || !power) return -1;
    hipLaunchKernelGGL(k_iq_power, dim3(1), dim3(64), 0, rf_raw_stream(s), td, n, power);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int rf_launch_iq_noise(const double *td, long long n, int noise, float std_f, uint64_t state0,
                                  uint64_t inc, float *out, rf_stream_t *s) {
    if (n <= 0) return 0;
    if (!td || !out || n >= (1LL << 40)) return -1;
    const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
    hipLaunchKernelGGL(k_iq_noise, dim3(blocks), dim3(256), 0, rf_raw_stream(s), td, n, noise, std_f, state0, inc,
                       out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
