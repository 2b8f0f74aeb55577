// rg_noalloc.hip -- rg_histogram_noalloc: the rho(T) histogram of device-resident samples with no
// allocation at all (no plan, no hipMalloc/hipFree, no host buffers), for callers that cannot
// allocate.  Its device scratch -- the bin edges and the 64-bit counters that k_histogram fills --
// is the file-scope array `scratch` below, one copy per GPU, held until the process ends.  Host code
// never uses the array's name as an address: its address on a device comes from hipGetSymbolAddress,
// taken once per device (with that device current) on the first call there and kept in a table.
// Every call on one GPU shares that GPU's copy, so calls that overlap on one GPU race; each call
// clears the counters it uses with a fill kernel before counting.
#include <mutex>

#include "rg_internal.h"

namespace {

constexpr int kNoallocMaxBins = 4096;  // RG_ENOTSUP above
constexpr size_t kBinsBytes = sizeof(double) * (kNoallocMaxBins + 1);
constexpr size_t kScratchBytes = kBinsBytes + sizeof(unsigned long long) * (kNoallocMaxBins + 2);  // 64 KiB
constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device` (which is current): looked up once per device.
rg_status scratch_address(int device, unsigned char **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return RG_EDEVICE;
        g_scratch_ptr[device] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch_ptr[device];
    return RG_OK;
}

__global__ __launch_bounds__(256) void k_zero_counts(unsigned long long *__restrict__ c, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) c[i] = 0ull;
}

struct CurrentDevice {
    int prev = -1;
    bool ok = false;
    explicit CurrentDevice(int dev)
    {
        ok = hipGetDevice(&prev) == hipSuccess && hipSetDevice(dev) == hipSuccess;
    }
    ~CurrentDevice()
    {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

}  // namespace

extern "C" rg_status rg_histogram_noalloc(int32_t device, void *stream, const double *samples_dev, int64_t nsamp,
                                          const double *bins, int32_t nbins, int64_t *counts, int64_t *missing,
                                          int64_t *dropped)
{
    if (!bins || !counts || nbins < 1 || nsamp < 0 || (!samples_dev && nsamp > 0)) return RG_EINVAL;
    if (nbins > kNoallocMaxBins) return RG_ENOTSUP;
    for (int i = 0; i < nbins; i++)
        if (!(bins[i] <= bins[i + 1])) return RG_EINVAL;  // check_histogram, src/base_utils.f90:468-482
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev || device >= kMaxDevices)
        return RG_EINVAL;
    CurrentDevice cur(device);
    if (!cur.ok) return RG_EDEVICE;
    unsigned char *base = nullptr;
    rg_status st = scratch_address(device, &base);
    if (st) return st;
    double *d_bins = reinterpret_cast<double *>(base);
    unsigned long long *d_cnt = reinterpret_cast<unsigned long long *>(base + kBinsBytes);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (hipMemcpyAsync(d_bins, bins, sizeof(double) * (nbins + 1), hipMemcpyHostToDevice, s) != hipSuccess)
        return RG_EDEVICE;
    hipLaunchKernelGGL(k_zero_counts, dim3(1), dim3(256), 0, s, d_cnt, nbins + 2);
    if (nsamp > 0) rg_launch_histogram_on(s, samples_dev, nsamp, d_bins, nbins, d_cnt);
    if (hipGetLastError() != hipSuccess) return RG_EDEVICE;
    // the counters are unsigned 64-bit counts far below 2^63: their bits are the int64 values
    int64_t tail[2] = {0, 0};
    if (hipMemcpyAsync(counts, d_cnt, sizeof(int64_t) * nbins, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(tail, d_cnt + nbins, sizeof(int64_t) * 2, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return RG_EDEVICE;
    if (missing) *missing = tail[0];
    if (dropped) *dropped = tail[1];
    return RG_OK;
}
