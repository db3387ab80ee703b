// spaqr_geqp3_batched_noalloc: the grouped geqp3 + choose_rank of spaqr_geqp3_batched for
// callers that cannot allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).  Its device scratch is a file-scope array of fixed size.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/spaqr_gpu.h"
#include "errors.hpp"
#include "kernels.hpp"
#include "rrqr.hpp"

using namespace spaqr;

// ---------------------------------------------------------------- geqp3 without allocation
// Device scratch of spaqr_geqp3_batched_noalloc: the cooperative exchange slots, barrier
// counters, error flag and descriptor tables of one call.  Every device the library runs on
// holds its own copy from module load until the process ends.  Host code never passes the
// array's name to the device: its address on the current device comes from
// hipGetSymbolAddress, once per device, and only that address is used.
constexpr size_t kGeqp3ScratchBytes = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[kGeqp3ScratchBytes];

namespace {
std::mutex g_scratch_mutex;
std::vector<void*> g_scratch_ptr;  // by device ordinal

void* geqp3_scratch() {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) throw std::runtime_error(std::string("hipGetDevice: ") + hipGetErrorString(e));
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

struct ScratchFull : std::runtime_error {
    using std::runtime_error::runtime_error;
};
}  // namespace

extern "C" {

int spaqr_geqp3_batched_noalloc(int prec, int count, const spaqr_geqp3_desc* d, double tol, double tie_rel,
                                void* stream) {
    if (prec != 32 && prec != 64) return -1;
    if (count < 0 || (count > 0 && !d)) return -2;
    for (int i = 0; i < count; ++i)
        if (d[i].k < 0 || d[i].w < 0) return -3;
    if (!(tol >= 0)) return -4;
    if (!(tie_rel >= 0 && tie_rel < 1)) return -5;
    if (count == 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    try {
        char* base = (char*)geqp3_scratch();
        size_t used = 0;
        std::vector<std::vector<char>> keep;  // host copies behind the queued uploads
        // Bump allocation inside the scratch array: geqp3_grouped makes every allocation and
        // upload before its first launch, so a call that does not fit stops before launching.
        RrqrAlloc mem{[&](size_t b) -> void* {
                          const size_t need = (b + 255) & ~(size_t)255;
                          if (used + need > kGeqp3ScratchBytes)
                              throw ScratchFull("geqp3 without allocation: the call needs more than the " +
                                                std::to_string(kGeqp3ScratchBytes >> 20) + " MiB of device scratch");
                          void* p = base + used;
                          used += need;
                          return p;
                      },
                      nullptr};
        mem.upload = [&](const void* h, size_t b) -> void* {
            void* p = mem.alloc(b);
            keep.emplace_back((const char*)h, (const char*)h + b);
            if (b && hipMemcpyAsync(p, keep.back().data(), b, hipMemcpyHostToDevice, s) != hipSuccess)
                throw std::runtime_error("geqp3 without allocation: descriptor upload failed");
            return p;
        };
        const int zero = 0;
        int* d_err = (int*)mem.upload(&zero, sizeof(int));  // cleared on every call
        if (prec == 64) {
            std::vector<Geqp3Task<double>> v((const Geqp3Task<double>*)d, (const Geqp3Task<double>*)d + count);
            geqp3_grouped<double>(v, tol, tie_rel, s, mem, d_err);
        } else {
            std::vector<Geqp3Task<float>> v((const Geqp3Task<float>*)d, (const Geqp3Task<float>*)d + count);
            geqp3_grouped<float>(v, tol, tie_rel, s, mem, d_err);
        }
        int herr = 0;
        hipError_t e = hipMemcpyAsync(&herr, d_err, sizeof(int), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) { api_error() = hipGetErrorString(e); return -100; }
        if (herr) { api_error() = "cooperative geqp3: barrier wait exceeded its spin limit"; return -100; }
        return 0;
    } catch (ScratchFull& ex) {
        api_error() = ex.what();
        (void)hipStreamSynchronize(s);  // the queued uploads into the scratch finish before it is reused
        return -6;
    } catch (std::exception& ex) {
        api_error() = ex.what();
        (void)hipStreamSynchronize(s);
        return -100;
    }
}

int spaqr_geqp3_noalloc_scratch_bytes(void) { return (int)kGeqp3ScratchBytes; }

}  // extern "C"
