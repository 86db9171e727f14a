// air_fbp.hip -- filtered back projection (auxil/fbp.m:31-80) on the device.
//
// The sinogram b (nr rays x na angles, column-major) is zero-padded per angle
// to order = max(64, 2^nextpow2(2 nr)) samples, transformed with a batched
// real-to-complex FFT (FP64), multiplied by the real ramp filter (optionally
// windowed), transformed back and cut to nr samples; the caller then
// back-projects with A'.  The reference filters the full complex spectrum with
// the symmetric extension [f; f(end-1:-1:2)] (:68-76); for a real sinogram and
// a real even filter the half spectrum carries the same information, so the
// device works on order/2 + 1 bins and the inverse transform returns the real
// part directly.
#include <cmath>
#include <vector>

#include "air_internal.h"

namespace air {

namespace {

__global__ void k_fbp_pad(const double* __restrict__ b, int64_t nr, int64_t na, int64_t order,
                          double* __restrict__ out) {
    const int64_t total = na * order;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = q / order, r = q - c * order;
        out[q] = r < nr ? b[c * nr + r] : 0.0;
    }
}

__global__ void k_fbp_filter(hipfftDoubleComplex* __restrict__ F, const double* __restrict__ f, int64_t nh,
                             int64_t na) {
    const int64_t total = na * nh;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
        const double g = f[q % nh];
        F[q].x *= g;
        F[q].y *= g;
    }
}

// cut to nr samples, undo the unnormalised inverse transform, narrow to FP32
__global__ void k_fbp_take(const double* __restrict__ in, int64_t nr, int64_t na, int64_t order,
                           float* __restrict__ out) {
    const int64_t total = na * nr;
    const double inv = 1.0 / (double)order;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = q / nr, r = q - c * nr;
        out[q] = (float)(in[c * order + r] * inv);
    }
}

unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 1 << 16)); }

struct FftPlan {
    hipfftHandle h = 0;
    bool made = false;
    ~FftPlan() {
        if (made) (void)hipfftDestroy(h);
    }
};

}  // namespace

// fbp.m:45-68: the first order/2 + 1 entries of the frequency response.
std::vector<double> fbp_filter_response(int64_t nr, int filter, int64_t* order_out) {
    int64_t pw = 1;
    while (pw < 2 * nr) pw <<= 1;                      // 2^nextpow2(2 nr)
    const int64_t order = std::max<int64_t>(64, pw);
    const int64_t nh = order / 2 + 1;
    std::vector<double> f(nh);
    const double nmax = (double)(order / 2);
    for (int64_t k = 0; k < nh; ++k) f[k] = (double)k / nmax;
    f[0] = 0.4 / (double)order;
    f[nh - 1] = 1 - f[0];
    for (int64_t k = 1; k < nh; ++k) {
        const double w = 2 * M_PI * (double)k / (double)order;
        switch (filter) {
        case AIR_FBP_RAMLAK: break;
        case AIR_FBP_SHEPPLOGAN: {                      // sinc(t) = sin(pi t)/(pi t), t = w/(2 pi)
            const double t = w / (2 * M_PI);
            f[k] = f[k] * (std::sin(M_PI * t) / (M_PI * t));
            break;
        }
        case AIR_FBP_COSINE: f[k] = f[k] * std::cos(w / 2); break;
        case AIR_FBP_HAMMING: f[k] = f[k] * (.54 + .46 * std::cos(w)); break;
        case AIR_FBP_HANN: f[k] = f[k] * (1 + std::cos(w)) / 2; break;
        default: throw Err{AIR_E_ARG, "Iinvalid filter"};
        }
    }
    *order_out = order;
    return f;
}

void fbp_filter_sinogram(const double* b_host, int64_t nr, int64_t na, int filter, float* out, hipStream_t s) {
    int64_t order = 0;
    const std::vector<double> f = fbp_filter_response(nr, filter, &order);
    AIR_CHECK(order < INT32_MAX && na < INT32_MAX, AIR_E_UNSUPPORTED, "fbp: sinogram too large for one transform");
    const int64_t nh = order / 2 + 1;
    DevBuf db, pad, spec, df;
    db.alloc(8 * (size_t)std::max<int64_t>(nr * na, 1));
    pad.alloc(8 * (size_t)std::max<int64_t>(na * order, 1));
    spec.alloc(16 * (size_t)std::max<int64_t>(na * nh, 1));
    df.alloc(8 * (size_t)nh);
    AIR_HIP(hipMemcpyAsync(db.p, b_host, 8 * (size_t)(nr * na), hipMemcpyHostToDevice, s));
    AIR_HIP(hipMemcpyAsync(df.p, f.data(), 8 * (size_t)nh, hipMemcpyHostToDevice, s));
    if (na == 0 || nr == 0) return;
    k_fbp_pad<<<grid_for(na * order), 256, 0, s>>>(db.as<double>(), nr, na, order, pad.as<double>());
    AIR_HIP(hipGetLastError());
    int n[1] = {(int)order};
    FftPlan fw, bw;
    AIR_CHECK(hipfftPlanMany(&fw.h, 1, n, nullptr, 1, (int)order, nullptr, 1, (int)nh, HIPFFT_D2Z, (int)na) ==
                  HIPFFT_SUCCESS,
              AIR_E_HIP, "hipfftPlanMany failed for the fbp forward transform");
    fw.made = true;
    AIR_CHECK(hipfftPlanMany(&bw.h, 1, n, nullptr, 1, (int)nh, nullptr, 1, (int)order, HIPFFT_Z2D, (int)na) ==
                  HIPFFT_SUCCESS,
              AIR_E_HIP, "hipfftPlanMany failed for the fbp inverse transform");
    bw.made = true;
    AIR_CHECK(hipfftSetStream(fw.h, s) == HIPFFT_SUCCESS && hipfftSetStream(bw.h, s) == HIPFFT_SUCCESS, AIR_E_HIP,
              "hipfftSetStream failed");
    hipfftDoubleComplex* F = (hipfftDoubleComplex*)spec.p;
    AIR_CHECK(hipfftExecD2Z(fw.h, pad.as<double>(), F) == HIPFFT_SUCCESS, AIR_E_HIP, "hipfftExecD2Z failed");
    k_fbp_filter<<<grid_for(na * nh), 256, 0, s>>>(F, df.as<double>(), nh, na);
    AIR_HIP(hipGetLastError());
    AIR_CHECK(hipfftExecZ2D(bw.h, F, pad.as<double>()) == HIPFFT_SUCCESS, AIR_E_HIP, "hipfftExecZ2D failed");
    k_fbp_take<<<grid_for(na * nr), 256, 0, s>>>(pad.as<double>(), nr, na, order, out);
    AIR_HIP(hipGetLastError());
    AIR_HIP(hipStreamSynchronize(s));
}

}  // namespace air
