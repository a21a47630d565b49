// Layer 1 of the C ABI: a stand-alone astrocyte population on the device,
// with the meaning of the reference's compile-time CUDA backend
// (include/astrosimgpu/astrocyte_cuda.hpp:13-27): the host arrays stay the
// canonical state, the handle owns device copies from create to destroy,
// destroy writes Ca, IP3 and h back, n == 0 or a null handle turns every
// call into a no-op, and a device failure ends the process (the interface
// has no error channel).
//
// Design here:
//  * One device allocation per handle. The eight per-cell arrays are planes
//    of one [8][stride] block of doubles, stride = n rounded up to 64 cells
//    (one wavefront reads a whole 512-byte line per plane).
//  * One HIP stream per handle (default priority). update() is queued on it
//    and returns at once; push, pull and destroy queue their copy on the
//    same stream and wait for it, so every transfer sees the updates queued
//    before it, and host buffers are free again when a call returns.
//  * The update is the whole-step engine's FP64 astrocyte math (cell_math.h:
//    the reference's operation order, glibc-exact exp/log/cos), one cell per
//    lane, in a grid sized to the cell count.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#include "asg.h"
#include "cell_math.h"
#include "rng.h"

namespace {

// Planes of the handle's block.
enum AstroPlane : int { kCa, kIP3, kH, kInput, kCaTot, kIP30, kTau, kDelta, kPlanes };

constexpr int kLanes = 256;  // four wavefronts per workgroup

[[noreturn]] void die(hipError_t err, const char* what, const char* file, int line) {
    std::fprintf(stderr, "asg_astro: %s failed with '%s' (%s:%d)\n", what,
                 hipGetErrorString(err), file, line);
    std::abort();
}

inline void ok(hipError_t err, const char* what, const char* file, int line) {
    if (err != hipSuccess) {
        die(err, what, file, line);
    }
}
#define ASTRO_OK(call) ok((call), #call, __FILE__, __LINE__)

struct AstroLaunch {
    asg::AstroK c;
    double h, half_h, h_sixth;
    int32_t substeps;
    int32_t per_cell_noise;  // 1: each cell draws its own normal
    double noise_std, shared_noise;
    uint64_t noise_key0;     // rng_key(noise_seed, noise_index * 1000003): cell i adds i * phi
};

// One step of every cell (astrocyte.cpp:233-329 / astrocyte_cuda.cu:27-58):
// noise current, IP3 jump, RK4 substeps and clamps, then the input plane is
// cleared for the next push.
__global__ __launch_bounds__(kLanes) void k_astro_layer1(double* __restrict__ block, size_t stride,
                                                          uint32_t n, AstroLaunch L) {
    const uint32_t i = blockIdx.x * kLanes + threadIdx.x;
    if (i >= n) {
        return;
    }
    double* __restrict__ Ca = block + kCa * stride;
    double* __restrict__ IP3 = block + kIP3 * stride;
    double* __restrict__ hh = block + kH * stride;
    double* __restrict__ input = block + kInput * stride;
    const double* __restrict__ Ca_tot = block + kCaTot * stride;
    const double* __restrict__ IP3_0 = block + kIP30 * stride;
    const double* __restrict__ tau = block + kTau * stride;
    const double* __restrict__ delta = block + kDelta * stride;
    double noise = L.shared_noise;
    if (L.per_cell_noise) {
        // stream noise_index * 1000003 + i: its key advances by phi per cell
        const uint64_t key = L.noise_key0 + static_cast<uint64_t>(i) * asg::kStreamMul;
        double u1 = asg::rng_uniform_keyed(key, 0);
        if (u1 <= 1e-300) {
            u1 = 1e-300;
        }
        const double u2 = asg::rng_uniform_keyed(key, 1);
        noise = L.noise_std * (asg::asg_sqrt(-2.0 * asg::asg_log(u1)) * asg::asg_cos(asg::kTwoPi * u2));
    }
    double ca = Ca[i], ip3 = IP3[i], h = hh[i];
    asg::astro_step(L.c, Ca_tot[i], IP3_0[i], tau[i], delta[i], input[i], noise, L.h, L.half_h,
                    L.h_sixth, L.substeps, ca, ip3, h);
    Ca[i] = ca;
    IP3[i] = ip3;
    hh[i] = h;
    input[i] = 0.0;
}

}  // namespace

struct asg_astro {
    uint32_t n = 0;
    size_t stride = 0;
    double* block = nullptr;
    hipStream_t stream = nullptr;

    double* plane(int k) const { return block + static_cast<size_t>(k) * stride; }
    size_t plane_bytes() const { return static_cast<size_t>(n) * sizeof(double); }

    void upload(int k, const double* host) const {
        ASTRO_OK(hipMemcpyAsync(plane(k), host, plane_bytes(), hipMemcpyHostToDevice, stream));
    }
    void download(int k, double* host) const {
        ASTRO_OK(hipMemcpyAsync(host, plane(k), plane_bytes(), hipMemcpyDeviceToHost, stream));
    }
    void wait() const { ASTRO_OK(hipStreamSynchronize(stream)); }
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
    s->stride = (static_cast<size_t>(n) + 63) & ~static_cast<size_t>(63);
    ASTRO_OK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    ASTRO_OK(hipMalloc(&s->block, kPlanes * s->stride * sizeof(double)));
    const double* init[kPlanes] = {Ca, IP3, h, nullptr, Ca_tot, IP3_0, tau_IP3, delta_IP3};
    for (int k = 0; k < kPlanes; ++k) {
        if (init[k] != nullptr) {
            s->upload(k, init[k]);
        }
    }
    ASTRO_OK(hipMemsetAsync(s->plane(kInput), 0, s->plane_bytes(), s->stream));
    s->wait();
    return s;
}

void asg_astro_destroy(asg_astro* s, double* Ca, double* IP3, double* h) {
    if (s == nullptr) {
        return;
    }
    if (s->n > 0) {
        s->download(kCa, Ca);
        s->download(kIP3, IP3);
        s->download(kH, h);
        s->wait();
        ASTRO_OK(hipFree(s->block));
        ASTRO_OK(hipStreamDestroy(s->stream));
    }
    delete s;
}

void asg_astro_push_input(asg_astro* s, const double* ip3_input) {
    if (s == nullptr || s->n == 0) {
        return;
    }
    s->upload(kInput, ip3_input);
    s->wait();
}

void asg_astro_pull_calcium(asg_astro* s, double* Ca) {
    if (s == nullptr || s->n == 0) {
        return;
    }
    s->download(kCa, Ca);
    s->wait();
}

void asg_astro_update(asg_astro* s, const asg_astro_constants* c, double h_step, int32_t substeps,
                      double noise_std, int32_t independent_noise, double shared_noise,
                      uint64_t noise_seed, uint64_t noise_index) {
    if (s == nullptr || s->n == 0) {
        return;
    }
    AstroLaunch L{};
    L.c = asg::make_astro_k(*c);
    L.h = h_step;
    L.half_h = 0.5 * h_step;
    L.h_sixth = h_step / 6.0;
    L.substeps = substeps;
    L.per_cell_noise = noise_std > 0.0 && independent_noise != 0;
    L.noise_std = noise_std;
    L.shared_noise = shared_noise;
    L.noise_key0 = asg::rng_key(noise_seed, noise_index * asg::kStreamStride);
    const uint32_t groups = (s->n + kLanes - 1) / kLanes;
    hipLaunchKernelGGL(k_astro_layer1, dim3(groups), dim3(kLanes), 0, s->stream, s->block,
                       s->stride, s->n, L);
    ASTRO_OK(hipGetLastError());
}

}  // extern "C"
