// FP64 diagnostic reductions that allocate nothing: the kernels read the cell values straight
// from the population storage (no per-cell double fields) and keep their per-block partials
// and the result in one device-resident array of fixed size, so that callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops) can take the
// diagnostics of Thermal_NusseltRBC / check_variables (MaterKit/inc/mkFluidsLBM.h:55,67-68).
//
// The array is file-scope device memory: every GPU holds its own copy at its own address
// until the process ends.  Its host-side name is not a device address, so the address is
// looked up once per device (noalloc_prepare, called by mkg_create with the handle's device
// current) and only that address is handed to kernels and copies.  Calls on one GPU share the
// array: two calls that overlap on one GPU race.  Each call writes every partial it reads
// (one per launched block) and the result slots before reading them, so nothing carries over
// from an earlier call.
#include <mutex>
#include <string>

#include "cellred.h"
#include "kernels.h"

namespace mkg {

// Largest grid served: 2^25 cells (e.g. 320^3), one partial block of 256 cells each.
constexpr uint32_t kNoAllocMaxCells = 1u << 25;
constexpr uint32_t kNoAllocMaxBlocks = kNoAllocMaxCells / 256u;
constexpr size_t kNoAllocOut = (size_t)kNoAllocMaxBlocks * kRedMax;  // offset of the result
constexpr size_t kNoAllocDoubles = kNoAllocOut + 16;                  // 4 MiB + 128 B

static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};
}  // namespace

// Device address of `scratch` on device dev (looked up once per device, with dev current).
int noalloc_prepare(int dev, std::string* err) {
    if (dev < 0 || dev >= kMaxDevices) {
        *err = "device index out of range for the reduction scratch table";
        return -1;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_ptr[dev]) return 0;
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) {
        *err = "hipGetDevice failed";
        return -1;
    }
    if (cur != dev && hipSetDevice(dev) != hipSuccess) {
        *err = "hipSetDevice failed";
        return -1;
    }
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !ptr) {
        *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return -1;
    }
    g_ptr[dev] = (double*)ptr;
    return 0;
}

// Cell values straight from the storage: MACRO 0 (phase field only), 1 (scalars only, no
// flow read), 2 (flow and scalars).
template <class R, int FQ, int MACRO>
__global__ void __launch_bounds__(256) k_reduce_direct(const StepArgs a, int which,
                                                       double* __restrict__ partials) {
    double v[kRedMax] = {0, 0, 0, 0};
    const bool is_max = reduce_is_max(which);
    if (is_max)
        for (int k = 0; k < kRedMax; k++) v[k] = -1e300;
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t < a.N && a.flag[t] != kIsWalls) {
        CellMacro m;
        if constexpr (MACRO > 0) {
            Cell c;
            locate_at(a, t, c);
            m = cell_macro<R, MACRO == 2 ? FQ : 0>(a, c, false, which == 2);
        }
        const R* ph = (const R*)a.phi_cur;
        const double phi = ph ? (double)ph[t] : 0.0;
        reduce_value(a, which, t, m.rho, m.u, m.T, m.C, phi,
                     [&](uint32_t k) { return (double)ph[k]; }, v);
    }
    block_reduce(v, kRedMax, is_max, partials + (size_t)blockIdx.x * kRedMax);
}

__global__ void __launch_bounds__(256) k_reduce_direct_final(const double* partials,
                                                             uint32_t nb, int is_max,
                                                             double* out) {
    reduce_partials(partials, nb, is_max != 0, out);
}

template <class R>
__global__ void __launch_bounds__(256) k_nonfinite_direct(const R* p, size_t n,
                                                          double* partials) {
    nonfinite_partial<R>(p, n, partials);
}

template <class R, int FQ>
static void launch_direct(const StepArgs& a, int which, int macro, double* part,
                          hipStream_t s) {
    dim3 g(a.nblk);
    if (macro == 0) k_reduce_direct<R, FQ, 0><<<g, 256, 0, s>>>(a, which, part);
    else if (macro == 1) k_reduce_direct<R, FQ, 1><<<g, 256, 0, s>>>(a, which, part);
    else k_reduce_direct<R, FQ, 2><<<g, 256, 0, s>>>(a, which, part);
}

template <class R>
static int launch_direct_fq(int fq, const StepArgs& a, int which, int macro, double* part,
                            hipStream_t s) {
    if (fq == 0) launch_direct<R, 0>(a, which, macro, part, s);
    else if (fq == 19) launch_direct<R, 19>(a, which, macro, part, s);
    else if (fq == 27) launch_direct<R, 27>(a, which, macro, part, s);
    else return -1;
    return 0;
}

// Reduction `which` into host[0..3] (before the host-side scaling of mkg_reduce).  pops and
// sizes list the population buffers for MKG_R_NONFINITE.  macro: see k_reduce_direct.
int reduce_noalloc(int dev, int fq, int fp32, const StepArgs& a, int which, int macro,
                   void* const* pops, const size_t* sizes, int npops, hipStream_t s,
                   double* host, std::string* err) {
    if (dev < 0 || dev >= kMaxDevices || !g_ptr[dev]) {
        *err = "reduction scratch not prepared on this device";
        return -1;
    }
    if (a.N > kNoAllocMaxCells || a.nblk > kNoAllocMaxBlocks) {
        *err = "grid of " + std::to_string(a.N) + " cells exceeds the " +
               std::to_string(kNoAllocMaxCells) +
               "-cell maximum of the device-resident reduction scratch";
        return -1;
    }
    double* part = g_ptr[dev];
    double* out = part + kNoAllocOut;
    auto fetch = [&](double* dst) -> int {
        hipError_t e = hipMemcpyAsync(dst, out, kRedMax * sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            *err = std::string("reduction read-back: ") + hipGetErrorString(e);
            return -1;
        }
        return 0;
    };
    if (which == 8) {  // non-finite populations, lattice by lattice
        const uint32_t nb = 1024;
        double total = 0;
        for (int k = 0; k < npops; k++) {
            if (fp32) k_nonfinite_direct<float><<<nb, 256, 0, s>>>((const float*)pops[k], sizes[k], part);
            else k_nonfinite_direct<double><<<nb, 256, 0, s>>>((const double*)pops[k], sizes[k], part);
            k_reduce_direct_final<<<1, 256, 0, s>>>(part, nb, 0, out);
            double v[kRedMax];
            if (fetch(v)) return -1;
            total += v[0];
        }
        host[0] = total;
        return 0;
    }
    const int rc = fp32 ? launch_direct_fq<float>(fq, a, which, macro, part, s)
                        : launch_direct_fq<double>(fq, a, which, macro, part, s);
    if (rc) {
        *err = "no reduction kernel for this flow lattice";
        return -1;
    }
    k_reduce_direct_final<<<1, 256, 0, s>>>(part, a.nblk, reduce_is_max(which) ? 1 : 0, out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        *err = std::string("reduction launch: ") + hipGetErrorString(e);
        return -1;
    }
    return fetch(host);
}

}  // namespace mkg
