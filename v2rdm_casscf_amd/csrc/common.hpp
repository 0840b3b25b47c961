// common.hpp -- host-side helpers shared by the C ABI translation units:
// HIP error checks, an owning device buffer, the error slot of
// v2g_last_error() and the try / catch wrapper of every entry point.
#pragma once
#include <hip/hip_runtime.h>

#include <chrono>
#include <stdexcept>
#include <string>
#include <vector>

namespace v2g {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw ::v2g::HipError(std::string(#expr) + ": " + hipGetErrorString(e_));       \
    } while (0)

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { if (p) (void)hipFree(p); }
    void alloc(size_t count)
    {
        if (p) { (void)hipFree(p); p = nullptr; }
        n = count;
        if (count) HIPCHK(hipMalloc(&p, sizeof(T) * count));
    }
    void upload(const std::vector<T> &v)
    {
        alloc(v.size());
        if (!v.empty()) HIPCHK(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    }
};

inline double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// the calling thread's last error message (v2g_last_error)
std::string &last_error();

}  // namespace v2g

// every extern "C" entry point: int status, 0 = success; -1 invalid argument,
// -2 HIP runtime error, -3 anything else; message in v2g_last_error()
#define API_BEGIN try {
#define API_END                                                                           \
    return 0;                                                                             \
    }                                                                                     \
    catch (const std::invalid_argument &e) { ::v2g::last_error() = e.what(); return -1; } \
    catch (const ::v2g::HipError &e) { ::v2g::last_error() = e.what(); return -2; }       \
    catch (const std::exception &e) { ::v2g::last_error() = e.what(); return -3; }
