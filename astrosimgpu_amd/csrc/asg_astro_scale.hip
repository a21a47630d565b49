// The isolated astrocyte update at scale: N Li-Rinzel cells in SoA arrays
// (Ca, IP3, h), one launch per time step that reads each cell's state and
// its presynaptic spike count from HBM and writes the state back -- the
// astrocyte half of the reference's step (astro_update_kernel,
// astrocyte_cuda.cu:27-58; astro_advance, astrocyte_kernel.hpp:71-103) with
// the IP3 jump delta_IP3 * count, measured as HBM traffic per step.
//
// Inputs are synthetic and reproducible from the seed alone:
//   initial state   Ca ~ U[0.05, 0.40], IP3 ~ U[0.01, 0.50], h ~ U[0.50, 0.95]
//                   from rng_uniform(seed ^ kSaltInit, i, 0 | 1 | 2);
//   per-cell        (randomised variant) Ca_tot, IP3_0, tau_IP3, delta_IP3 drawn
//   parameters      around their nominal values like the reference's
//                   draw_scaled (astrocyte.cpp:13-23): normal with sd
//                   var * |mean|, redrawn until inside [lower, upper] * mean,
//                   from stream 4 i + p of seed ^ kSaltParams;
//   spike counts    Poisson(lambda) per cell and step from
//                   rng_poisson_fresh(seed, step * N + i), written to an int32
//                   array in HBM by a separate launch before each update.
// Both precisions start from the same FP64 initial state (rounded to FP32)
// and read the same counts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "asg.h"
#include "cell_math.h"
#include "cell_math_f32.h"
#include "rng.h"

extern "C" asg_status asg_host_fail(asg_status code, const char* what);

namespace {

constexpr int kBlock = 256;
constexpr uint64_t kSaltInit = 0x6A09E667F3BCC909ULL;
constexpr uint64_t kSaltParams = 0xBB67AE8584CAA73BULL;

struct HipFail {
    std::string what;
};

void check(hipError_t e, const char* call) {
    if (e != hipSuccess) {
        throw HipFail{std::string(call) + ": " + hipGetErrorString(e)};
    }
}
#define SCALE_CHECK(call) check((call), #call)

__device__ double draw_scaled(uint64_t seed, uint64_t stream, double mean, double lower,
                              double upper, double var) {
    const double lo = mean < 0.0 ? mean * upper : mean * lower;
    const double hi = mean < 0.0 ? mean * lower : mean * upper;
    const double sd = var * fabs(mean);
    uint64_t ctr = 0;
    for (int tries = 0; tries < 256; ++tries) {
        const double v = mean + sd * asg::rng_normal_fresh(seed, stream, ctr);
        if (v >= lo && v <= hi) {
            return v;
        }
    }
    return mean;
}

__global__ __launch_bounds__(kBlock) void k_scale_init(uint32_t n, uint64_t seed, int32_t randomized,
                                                       asg_astro_scaling_params p, double* Ca,
                                                       double* IP3, double* h, double* Ca_tot,
                                                       double* IP3_0, double* tau, double* delta) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) {
        return;
    }
    const uint64_t s0 = seed ^ kSaltInit;
    Ca[i] = 0.05 + 0.35 * asg::rng_uniform(s0, i, 0);
    IP3[i] = 0.01 + 0.49 * asg::rng_uniform(s0, i, 1);
    h[i] = 0.50 + 0.45 * asg::rng_uniform(s0, i, 2);
    if (randomized) {
        const uint64_t s1 = seed ^ kSaltParams;
        const uint64_t b = 4 * static_cast<uint64_t>(i);
        Ca_tot[i] = draw_scaled(s1, b + 0, p.Ca_tot, p.lower, p.upper, p.var);
        IP3_0[i] = draw_scaled(s1, b + 1, p.IP3_0, p.lower, p.upper, p.var);
        tau[i] = draw_scaled(s1, b + 2, p.tau_IP3, p.lower, p.upper, p.var);
        delta[i] = draw_scaled(s1, b + 3, p.delta_IP3, p.lower, p.upper, p.var);
    }
}

__global__ __launch_bounds__(kBlock) void k_scale_counts(uint32_t n, uint64_t seed, int64_t step,
                                                         double lambda, double limit,
                                                         int32_t* counts) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        counts[i] = asg::rng_poisson_fresh(seed, static_cast<uint64_t>(step) * n + i, lambda,
                                           limit);
    }
}

__global__ __launch_bounds__(kBlock) void k_scale_to_f32(uint32_t n, const double* a, float* b) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        b[i] = static_cast<float>(a[i]);
    }
}

__global__ __launch_bounds__(kBlock) void k_scale_inv_f32(uint32_t n, const double* a, float* b) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        b[i] = static_cast<float>(1.0 / a[i]);
    }
}

// One step of every cell, FP64: the reference's astro_advance term for term.
__global__ __launch_bounds__(kBlock) void k_scale_update64(
    uint32_t n, asg::AstroK k, double hs, double half_h, double h_sixth, const int32_t* counts,
    double* __restrict__ Ca, double* __restrict__ IP3, double* __restrict__ h,
    const double* __restrict__ Ca_tot_c, const double* __restrict__ IP3_0_c,
    const double* __restrict__ tau_c, const double* __restrict__ delta_c, double Ca_tot,
    double IP3_0, double tau, double delta) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) {
        return;
    }
    double c = Ca[i], p = IP3[i], g = h[i];
    if (Ca_tot_c != nullptr) {
        Ca_tot = Ca_tot_c[i];
        IP3_0 = IP3_0_c[i];
        tau = tau_c[i];
        delta = delta_c[i];
    }
    asg::astro_step(k, Ca_tot, IP3_0, tau, delta, static_cast<double>(counts[i]), 0.0, hs,
                    half_h, h_sixth, 1, c, p, g);
    Ca[i] = c;
    IP3[i] = p;
    h[i] = g;
}

// One step of every cell, FP32 (state, constants and reciprocals in FP32).
__global__ __launch_bounds__(kBlock) void k_scale_update32(
    uint32_t n, asg::AstroKf k, float hs, float half_h, float h_sixth, const int32_t* counts,
    float* __restrict__ Ca, float* __restrict__ IP3, float* __restrict__ h,
    const float* __restrict__ Ca_tot_c, const float* __restrict__ IP3_0_c,
    const float* __restrict__ inv_tau_c, const float* __restrict__ delta_c, float Ca_tot,
    float IP3_0, float inv_tau, float delta) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) {
        return;
    }
    float c = Ca[i], p = IP3[i], g = h[i];
    if (Ca_tot_c != nullptr) {
        Ca_tot = Ca_tot_c[i];
        IP3_0 = IP3_0_c[i];
        inv_tau = inv_tau_c[i];
        delta = delta_c[i];
    }
    asg::astro_step_f32(k, Ca_tot, IP3_0, inv_tau, delta, static_cast<float>(counts[i]), 0.0f,
                        hs, half_h, h_sixth, 1, c, p, g);
    Ca[i] = c;
    IP3[i] = p;
    h[i] = g;
}

struct Buffers {
    std::vector<void*> ptrs;
    template <typename T>
    T* alloc(size_t count) {
        void* p = nullptr;
        SCALE_CHECK(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
        ptrs.push_back(p);
        return static_cast<T*>(p);
    }
    ~Buffers() {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

}  // namespace

extern "C" {

int64_t asg_astro_scaling_bytes(uint32_t n, int32_t precision, int32_t randomized) {
    const int64_t s = precision == ASG_FP32 ? 4 : 8;
    return static_cast<int64_t>(n) * (6 * s + 4 + (randomized ? 4 * s : 0));
}

asg_status asg_astro_scaling(uint32_t n, int32_t steps, int32_t precision,
                             const asg_astro_constants* constants,
                             const asg_astro_scaling_params* params, uint64_t seed,
                             double* Ca_out, double* IP3_out, double* h_out,
                             asg_astro_scaling_result* out) {
    if (n == 0 || steps < 0 || constants == nullptr || params == nullptr || out == nullptr ||
        (precision != ASG_FP64 && precision != ASG_FP32) || params->dt <= 0.0 ||
        params->lambda < 0.0 || params->lambda > 30.0) {
        return asg_host_fail(ASG_ERR_INVALID, "bad argument");
    }
    try {
        const asg_astro_scaling_params& p = *params;
        const bool f32 = precision == ASG_FP32;
        const int grid = static_cast<int>((n + kBlock - 1) / kBlock);
        hipStream_t st = nullptr;
        SCALE_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamDestroy(s); }
        } guard{st};
        Buffers b;
        double* Ca = b.alloc<double>(n);
        double* IP3 = b.alloc<double>(n);
        double* h = b.alloc<double>(n);
        int32_t* counts = b.alloc<int32_t>(n);
        double *Ca_tot = nullptr, *IP3_0 = nullptr, *tau = nullptr, *delta = nullptr;
        if (p.randomized) {
            Ca_tot = b.alloc<double>(n);
            IP3_0 = b.alloc<double>(n);
            tau = b.alloc<double>(n);
            delta = b.alloc<double>(n);
        }
        hipLaunchKernelGGL(k_scale_init, dim3(grid), dim3(kBlock), 0, st, n, seed, p.randomized,
                           p, Ca, IP3, h, Ca_tot, IP3_0, tau, delta);
        float *fCa = nullptr, *fIP3 = nullptr, *fh = nullptr;
        float *fCa_tot = nullptr, *fIP3_0 = nullptr, *finv_tau = nullptr, *fdelta = nullptr;
        if (f32) {
            fCa = b.alloc<float>(n);
            fIP3 = b.alloc<float>(n);
            fh = b.alloc<float>(n);
            hipLaunchKernelGGL(k_scale_to_f32, dim3(grid), dim3(kBlock), 0, st, n, Ca, fCa);
            hipLaunchKernelGGL(k_scale_to_f32, dim3(grid), dim3(kBlock), 0, st, n, IP3, fIP3);
            hipLaunchKernelGGL(k_scale_to_f32, dim3(grid), dim3(kBlock), 0, st, n, h, fh);
            if (p.randomized) {
                fCa_tot = b.alloc<float>(n);
                fIP3_0 = b.alloc<float>(n);
                finv_tau = b.alloc<float>(n);
                fdelta = b.alloc<float>(n);
                hipLaunchKernelGGL(k_scale_to_f32, dim3(grid), dim3(kBlock), 0, st, n, Ca_tot, fCa_tot);
                hipLaunchKernelGGL(k_scale_to_f32, dim3(grid), dim3(kBlock), 0, st, n, IP3_0, fIP3_0);
                hipLaunchKernelGGL(k_scale_inv_f32, dim3(grid), dim3(kBlock), 0, st, n, tau, finv_tau);
                hipLaunchKernelGGL(k_scale_to_f32, dim3(grid), dim3(kBlock), 0, st, n, delta, fdelta);
            }
        }
        SCALE_CHECK(hipGetLastError());
        const double hs = p.dt, half_h = 0.5 * p.dt, h_sixth = p.dt / 6.0;
        const double limit = asg::asg_exp(-p.lambda);
        const asg::AstroK k64 = asg::make_astro_k(*constants);
        const asg::AstroKf k32 = asg::make_astro_kf(*constants);
        std::vector<hipEvent_t> ev(2 * static_cast<size_t>(steps) + 2);
        for (hipEvent_t& e : ev) SCALE_CHECK(hipEventCreate(&e));
        struct EventGuard {
            std::vector<hipEvent_t>& v;
            ~EventGuard() {
                for (hipEvent_t e : v) (void)hipEventDestroy(e);
            }
        } eguard{ev};
        for (int32_t s = 0; s < steps; ++s) {
            hipLaunchKernelGGL(k_scale_counts, dim3(grid), dim3(kBlock), 0, st, n, seed,
                               static_cast<int64_t>(s), p.lambda, limit, counts);
            SCALE_CHECK(hipEventRecord(ev[2 * s], st));
            if (f32) {
                hipLaunchKernelGGL(k_scale_update32, dim3(grid), dim3(kBlock), 0, st, n, k32,
                                   static_cast<float>(hs), static_cast<float>(half_h),
                                   static_cast<float>(h_sixth), counts, fCa, fIP3, fh, fCa_tot,
                                   fIP3_0, finv_tau, fdelta, static_cast<float>(p.Ca_tot),
                                   static_cast<float>(p.IP3_0),
                                   static_cast<float>(1.0 / p.tau_IP3),
                                   static_cast<float>(p.delta_IP3));
            } else {
                hipLaunchKernelGGL(k_scale_update64, dim3(grid), dim3(kBlock), 0, st, n, k64, hs,
                                   half_h, h_sixth, counts, Ca, IP3, h, Ca_tot, IP3_0, tau,
                                   delta, p.Ca_tot, p.IP3_0, p.tau_IP3, p.delta_IP3);
            }
            SCALE_CHECK(hipEventRecord(ev[2 * s + 1], st));
        }
        SCALE_CHECK(hipGetLastError());
        SCALE_CHECK(hipStreamSynchronize(st));
        double update_ms = 0.0, total_ms = 0.0;
        for (int32_t s = 0; s < steps; ++s) {
            float ms = 0.0f;
            SCALE_CHECK(hipEventElapsedTime(&ms, ev[2 * s], ev[2 * s + 1]));
            update_ms += ms;
        }
        if (steps > 0) {
            float ms = 0.0f;
            SCALE_CHECK(hipEventElapsedTime(&ms, ev[0], ev[2 * (steps - 1) + 1]));
            total_ms = ms;
        }
        if (f32) {
            std::vector<float> tmp(n);
            auto back = [&](const float* src, double* dst) {
                if (dst == nullptr) return;
                SCALE_CHECK(hipMemcpy(tmp.data(), src, n * sizeof(float), hipMemcpyDeviceToHost));
                for (uint32_t i = 0; i < n; ++i) dst[i] = tmp[i];
            };
            back(fCa, Ca_out);
            back(fIP3, IP3_out);
            back(fh, h_out);
        } else {
            if (Ca_out) SCALE_CHECK(hipMemcpy(Ca_out, Ca, n * sizeof(double), hipMemcpyDeviceToHost));
            if (IP3_out) SCALE_CHECK(hipMemcpy(IP3_out, IP3, n * sizeof(double), hipMemcpyDeviceToHost));
            if (h_out) SCALE_CHECK(hipMemcpy(h_out, h, n * sizeof(double), hipMemcpyDeviceToHost));
        }
        out->update_ms = update_ms;
        out->total_ms = total_ms;
        out->bytes_per_step = static_cast<double>(asg_astro_scaling_bytes(n, precision, p.randomized));
        out->steps = steps;
        return ASG_OK;
    } catch (const HipFail& e) {
        return asg_host_fail(ASG_ERR_DEVICE, e.what.c_str());
    } catch (const std::bad_alloc&) {
        return asg_host_fail(ASG_ERR_NOMEM, "host allocation failed");
    }
}

}  // extern "C"
