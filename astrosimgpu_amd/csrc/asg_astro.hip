// Drop-in replacement of the reference's native CUDA astrocyte backend
// (include/astrosimgpu/astrocyte_cuda.hpp:13-27, src/astrocyte_cuda.cu).
// Same handle lifetime, same blocking transfers, same abort on any device
// error; the update kernel is the FP64 cell math the whole-step engine uses.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#include "asg.h"
#include "cell_math.h"
#include "rng.h"

namespace {

void check(hipError_t err, const char* what, const char* file, int line) {
    if (err != hipSuccess) {
        std::fprintf(stderr, "HIP error at %s:%d during %s: %s\n", file, line, what,
                     hipGetErrorString(err));
        std::abort();
    }
}
#define ASTRO_CHECK(call) check((call), #call, __FILE__, __LINE__)

constexpr int kBlock = 256;

// astro_update_kernel, astrocyte_cuda.cu:27-58: one cell per lane.
__global__ __launch_bounds__(kBlock) void k_astro_update(
    uint32_t n, asg::AstroK c, double h_step, double half_h, double h_sixth, int substeps,
    double noise_std, int independent_noise, double shared_noise, uint64_t noise_seed,
    uint64_t noise_index, double* __restrict__ Ca, double* __restrict__ IP3,
    double* __restrict__ h, double* __restrict__ ip3_input, const double* __restrict__ Ca_tot,
    const double* __restrict__ IP3_0, const double* __restrict__ tau_IP3,
    const double* __restrict__ delta_IP3) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) {
        return;
    }
    const double noise =
        (noise_std > 0.0 && independent_noise)
            ? noise_std * asg::rng_normal_clamped(noise_seed, noise_index * 1000003ULL + i)
            : shared_noise;
    double ca = Ca[i], ip3 = IP3[i], hh = h[i];
    asg::astro_step(c, Ca_tot[i], IP3_0[i], tau_IP3[i], delta_IP3[i], ip3_input[i], noise, h_step,
                    half_h, h_sixth, substeps, ca, ip3, hh);
    Ca[i] = ca;
    IP3[i] = ip3;
    h[i] = hh;
    ip3_input[i] = 0.0;
}

double* device_copy(const double* host, uint32_t n) {
    double* d = nullptr;
    ASTRO_CHECK(hipMalloc(&d, static_cast<size_t>(n) * sizeof(double)));
    ASTRO_CHECK(hipMemcpy(d, host, static_cast<size_t>(n) * sizeof(double), hipMemcpyHostToDevice));
    return d;
}

}  // namespace

struct asg_astro {
    uint32_t n = 0;
    double* Ca = nullptr;
    double* IP3 = nullptr;
    double* h = nullptr;
    double* ip3_input = nullptr;
    double* Ca_tot = nullptr;
    double* IP3_0 = nullptr;
    double* tau_IP3 = nullptr;
    double* delta_IP3 = nullptr;
};

extern "C" {

asg_astro* asg_astro_create(uint32_t n, const double* Ca, const double* IP3, const double* h,
                            const double* Ca_tot, const double* IP3_0, const double* tau_IP3,
                            const double* delta_IP3) {
    auto* s = new asg_astro;
    s->n = n;
    if (n == 0) {
        return s;
    }
    s->Ca = device_copy(Ca, n);
    s->IP3 = device_copy(IP3, n);
    s->h = device_copy(h, n);
    s->Ca_tot = device_copy(Ca_tot, n);
    s->IP3_0 = device_copy(IP3_0, n);
    s->tau_IP3 = device_copy(tau_IP3, n);
    s->delta_IP3 = device_copy(delta_IP3, n);
    ASTRO_CHECK(hipMalloc(&s->ip3_input, static_cast<size_t>(n) * sizeof(double)));
    ASTRO_CHECK(hipMemset(s->ip3_input, 0, static_cast<size_t>(n) * sizeof(double)));
    return s;
}

void asg_astro_destroy(asg_astro* s, double* Ca, double* IP3, double* h) {
    if (s == nullptr) {
        return;
    }
    if (s->n > 0) {
        const size_t bytes = static_cast<size_t>(s->n) * sizeof(double);
        ASTRO_CHECK(hipMemcpy(Ca, s->Ca, bytes, hipMemcpyDeviceToHost));
        ASTRO_CHECK(hipMemcpy(IP3, s->IP3, bytes, hipMemcpyDeviceToHost));
        ASTRO_CHECK(hipMemcpy(h, s->h, bytes, hipMemcpyDeviceToHost));
        for (double* p : {s->Ca, s->IP3, s->h, s->ip3_input, s->Ca_tot, s->IP3_0, s->tau_IP3,
                          s->delta_IP3}) {
            ASTRO_CHECK(hipFree(p));
        }
    }
    delete s;
}

void asg_astro_push_input(asg_astro* s, const double* ip3_input) {
    if (s == nullptr || s->n == 0) {
        return;
    }
    ASTRO_CHECK(hipMemcpy(s->ip3_input, ip3_input, static_cast<size_t>(s->n) * sizeof(double),
                          hipMemcpyHostToDevice));
}

void asg_astro_pull_calcium(asg_astro* s, double* Ca) {
    if (s == nullptr || s->n == 0) {
        return;
    }
    ASTRO_CHECK(hipMemcpy(Ca, s->Ca, static_cast<size_t>(s->n) * sizeof(double),
                          hipMemcpyDeviceToHost));
}

void asg_astro_update(asg_astro* s, const asg_astro_constants* c, double h_step, int32_t substeps,
                      double noise_std, int32_t independent_noise, double shared_noise,
                      uint64_t noise_seed, uint64_t noise_index) {
    if (s == nullptr || s->n == 0) {
        return;
    }
    const asg::AstroK k = asg::make_astro_k(*c);
    const int grid = static_cast<int>((s->n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_astro_update, dim3(grid), dim3(kBlock), 0, 0, s->n, k, h_step,
                       0.5 * h_step, h_step / 6.0, substeps, noise_std, independent_noise,
                       shared_noise, noise_seed, noise_index, s->Ca, s->IP3, s->h, s->ip3_input,
                       s->Ca_tot, s->IP3_0, s->tau_IP3, s->delta_IP3);
    ASTRO_CHECK(hipGetLastError());
}

}  // extern "C"
