// Error plumbing and small utility entry points.
#include "pf_common.hpp"

namespace pf {

static thread_local char g_last_error[512] = "";

int set_error(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

__global__ void k_fill_u32(uint32_t* __restrict__ dst, int64_t n, uint32_t v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = v;
}

int fill_u32(void* dst, int64_t n, uint32_t v, hipStream_t s) {
    if (n <= 0) return PF_OK;
    const int64_t nb = n / 256 + 1;
    const unsigned grid = (unsigned)(nb < 65536 ? nb : 65536);
    k_fill_u32<<<grid, 256, 0, s>>>(static_cast<uint32_t*>(dst), n, v);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

}  // namespace pf

PF_API const char* pf_last_error(void) { return pf::g_last_error; }

PF_API int pf_abi_version(void) { return 1; }

PF_API int pf_fill_u32(void* dst, int64_t n, uint32_t value, pf_stream_t stream) {
    PF_REQUIRE(dst != nullptr || n == 0, "pf_fill_u32: null destination");
    return pf::fill_u32(dst, n, value, pf::S(stream));
}

PF_API int pf_read_scalars(const void* dev_src, void* host_dst, size_t bytes, pf_stream_t stream) {
    PF_REQUIRE(dev_src && host_dst, "pf_read_scalars: null pointer");
    PF_CHECK_HIP(hipMemcpyAsync(host_dst, dev_src, bytes, hipMemcpyDeviceToHost, pf::S(stream)));
    PF_CHECK_HIP(hipStreamSynchronize(pf::S(stream)));
    return PF_OK;
}
