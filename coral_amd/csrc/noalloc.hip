// Batched banded LU factor and solve that allocates nothing: the work space of
// coral_amd_banded_solve (factor records and pivots) lives in one fixed
// file-scope device array per GPU, and the caller's buffers are device memory.
//
// `scratch` below is a host-side stand-in in host code, not a device address:
// its device address is looked up once per device with hipGetSymbolAddress
// (with that device current) and only that pointer is handed to kernels.
#include <algorithm>
#include <mutex>
#include <string>

#include "../../include/coral_amd.h"
#include "kernels.hpp"

namespace coral {
void set_last_error(const std::string &msg);
}

namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(CORAL_AMD_NOALLOC_SCRATCH_BYTES);
constexpr int MAX_DEVICES = 64;

static_assert(NOALLOC_SCRATCH_BYTES % 256 == 0, "scratch size");
__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_mu;
void *g_ptr[MAX_DEVICES] = {};   // device address of `scratch` per device (null until looked up)

#define NA_OK(call)                                                                          \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// the scratch array's address on `device`, which must be current
unsigned char *scratch_on(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[device]) {
        void *p = nullptr;
        NA_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw std::runtime_error("no device address for the scratch array");
        g_ptr[device] = p;
    }
    return static_cast<unsigned char *>(g_ptr[device]);
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

extern "C" int coral_amd_banded_solve_noalloc(int32_t device, int32_t n, int32_t kl, int32_t ku, int32_t nsys,
                                              double *ab, double *x, int32_t *info, void *stream) {
    try {
        if (!ab || !x || !info) throw std::runtime_error("null argument");
        if (n < 1 || kl < 0 || ku < 0 || nsys < 1 || kl > 255) throw std::runtime_error("bad band dimensions");
        if (device < 0 || device >= MAX_DEVICES) throw std::runtime_error("bad device index");
        const size_t kv = size_t(kl) + ku;
        const size_t need = coral_amd_banded_solve_noalloc_bytes(n, kl, ku, nsys);
        if (need > NOALLOC_SCRATCH_BYTES)
            throw std::runtime_error("banded_solve_noalloc: " + std::to_string(nsys) + " systems of n = " +
                                     std::to_string(n) + ", kl = " + std::to_string(kl) + ", ku = " +
                                     std::to_string(ku) + " need " + std::to_string(need) +
                                     " bytes of work space; the fixed array holds " +
                                     std::to_string(NOALLOC_SCRATCH_BYTES));
        NA_OK(hipSetDevice(device));
        unsigned char *base = scratch_on(device);
        // layout: multipliers [nsys][n][max(kl,1)], U records [nsys][n][kv+1], pivots [nsys][n]
        const size_t lbytes = align256(size_t(nsys) * n * std::max<size_t>(kl, 1) * sizeof(double2));
        const size_t ubytes = align256(size_t(nsys) * n * (kv + 1) * sizeof(double2));
        double2 *Lf = reinterpret_cast<double2 *>(base);
        double2 *Uf = reinterpret_cast<double2 *>(base + lbytes);
        unsigned char *piv = base + lbytes + ubytes;
        hipStream_t st = static_cast<hipStream_t>(stream);
        // every entry of Lf, Uf and piv is written by the factorization before
        // the solve reads it, so nothing is cleared
        coral::launch_factor_lapack(st, reinterpret_cast<double2 *>(ab), n, kl, ku, nsys, Lf, Uf, piv, info);
        coral::launch_solve(st, Lf, Uf, piv, n, kl, int(kv), nsys, reinterpret_cast<const double2 *>(x),
                            reinterpret_cast<double2 *>(x));
        return 0;
    } catch (const std::exception &e) {
        coral::set_last_error(e.what());
        return -1;
    }
}

extern "C" size_t coral_amd_banded_solve_noalloc_bytes(int32_t n, int32_t kl, int32_t ku, int32_t nsys) {
    if (n < 1 || kl < 0 || ku < 0 || nsys < 1) return 0;
    const size_t kv = size_t(kl) + ku;
    return align256(size_t(nsys) * n * std::max<size_t>(kl, 1) * sizeof(double2)) +
           align256(size_t(nsys) * n * (kv + 1) * sizeof(double2)) + align256(size_t(nsys) * n);
}
