// hip_proxy: device selection, library stream, memory operations and timers.
// Reference counterpart: src/cuda_proxy.cu:13-128 (device selection and CUDA
// runtime wrappers).  Design differences: one library stream per device,
// status codes instead of exit(), a fill kernel instead of memset so that the
// path stays capturable in a HIP graph.
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "hip_common.h"

namespace ca3 {

namespace {
constexpr int kMaxDev = 16;
hipStream_t g_stream[kMaxDev] = {};
hipEvent_t g_ev0[kMaxDev] = {}, g_ev1[kMaxDev] = {};
int g_cus[kMaxDev] = {};
std::mutex g_mu;
bool g_sync = true;
thread_local char g_err[512] = "";

int cur_dev()
{
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess) return 0;
    return (d >= 0 && d < kMaxDev) ? d : 0;
}
}  // namespace

hipStream_t stream()
{
    int d = cur_dev();
    if (!g_stream[d]) {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_stream[d]) {
            hipStream_t s = nullptr;
            if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess) g_stream[d] = s;
        }
    }
    return g_stream[d];
}

bool sync_mode() { return g_sync; }

int fail(hipError_t e, const char *what)
{
    snprintf(g_err, sizeof(g_err), "%s: %s", what, hipGetErrorString(e));
    fprintf(stderr, "[ca3dmm_amd] HIP error in %s\n", g_err);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return HIP_PROXY_ERR_NOMEM;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return HIP_PROXY_ERR_NODEV;
    return HIP_PROXY_ERR_RUNTIME;
}

int fail_msg(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    fprintf(stderr, "[ca3dmm_amd] %s\n", g_err);
    return code;
}

int finish()
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(e, "kernel launch");
    if (g_sync) {
        e = hipStreamSynchronize(stream());
        if (e != hipSuccess) return fail(e, "hipStreamSynchronize");
    }
    return HIP_PROXY_OK;
}

int cu_count()
{
    int d = cur_dev();
    if (!g_cus[d]) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, d) != hipSuccess || v <= 0)
            v = 256;
        g_cus[d] = v;
    }
    return g_cus[d];
}

}  // namespace ca3

namespace {

// Byte fill, 16-byte stores for the aligned body.
__global__ void fill_bytes_kernel(unsigned char *p, size_t head, size_t nvec, size_t tail,
                                  unsigned int word)
{
    const size_t tid = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
    const size_t stride = (size_t) gridDim.x * blockDim.x;
    const unsigned char b = (unsigned char) (word & 0xff);
    for (size_t i = tid; i < head; i += stride) p[i] = b;
    uint4 *v = reinterpret_cast<uint4 *>(p + head);
    const uint4 w = make_uint4(word, word, word, word);
    for (size_t i = tid; i < nvec; i += stride) v[i] = w;
    unsigned char *t = p + head + nvec * 16;
    for (size_t i = tid; i < tail; i += stride) t[i] = b;
}

}  // namespace

extern "C" {

int hip_proxy_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int select_hip_device_by_mpi_local_rank(void)
{
    static const char *keys[] = {"MPI_LOCALRANKID", "MV2_COMM_WORLD_LOCAL_RANK",
                                 "OMPI_COMM_WORLD_NODE_RANK", "SLURM_LOCALID", "PBS_O_VNODENUM"};
    int local_rank = 0;
    for (const char *k : keys) {
        const char *v = getenv(k);
        if (v) { local_rank = atoi(v); break; }
    }
    int n = hip_proxy_device_count();
    if (n <= 0) return ca3::fail_msg(HIP_PROXY_ERR_NODEV, "no HIP device visible");
    return hip_set_rt_dev_id(local_rank % n);
}

int hip_set_rt_dev_id(const int dev_id)
{
    CA3_CHECK(hipSetDevice(dev_id));
    return HIP_PROXY_OK;
}

int hip_proxy_set_sync(int sync)
{
    ca3::g_sync = (sync != 0);
    return HIP_PROXY_OK;
}

int hip_device_sync(void)
{
    CA3_CHECK(hipDeviceSynchronize());
    return HIP_PROXY_OK;
}

int hip_stream_sync(void *stream)
{
    hipStream_t s = stream ? (hipStream_t) stream : ca3::stream();
    CA3_CHECK(hipStreamSynchronize(s));
    return HIP_PROXY_OK;
}

void *hip_proxy_stream(void) { return (void *) ca3::stream(); }

int hip_timer_start(void)
{
    int d = ca3::cur_dev();
    if (!ca3::g_ev0[d]) {
        CA3_CHECK(hipEventCreate(&ca3::g_ev0[d]));
        CA3_CHECK(hipEventCreate(&ca3::g_ev1[d]));
    }
    CA3_CHECK(hipEventRecord(ca3::g_ev0[d], ca3::stream()));
    return HIP_PROXY_OK;
}

int hip_timer_stop(double *ms)
{
    int d = ca3::cur_dev();
    if (!ca3::g_ev0[d]) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "hip_timer_stop without start");
    CA3_CHECK(hipEventRecord(ca3::g_ev1[d], ca3::stream()));
    CA3_CHECK(hipEventSynchronize(ca3::g_ev1[d]));
    float f = 0.f;
    CA3_CHECK(hipEventElapsedTime(&f, ca3::g_ev0[d], ca3::g_ev1[d]));
    if (ms) *ms = (double) f;
    return HIP_PROXY_OK;
}

const char *hip_proxy_last_error(void) { return ca3::g_err; }

int hip_malloc_dev(void **dptr_, const size_t bytes)
{
    if (!dptr_) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "hip_malloc_dev: null out pointer");
    *dptr_ = nullptr;
    if (bytes == 0) return HIP_PROXY_OK;
    CA3_CHECK(hipMalloc(dptr_, bytes));
    return HIP_PROXY_OK;
}

int hip_malloc_host(void **hptr_, const size_t bytes)
{
    if (!hptr_) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "hip_malloc_host: null out pointer");
    *hptr_ = nullptr;
    if (bytes == 0) return HIP_PROXY_OK;
    CA3_CHECK(hipHostMalloc(hptr_, bytes, hipHostMallocDefault));
    return HIP_PROXY_OK;
}

int hip_memset_dev(void *dptr, const int value, const size_t bytes)
{
    if (bytes == 0) return HIP_PROXY_OK;
    if (!dptr) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "hip_memset_dev: null pointer");
    uintptr_t a = (uintptr_t) dptr;
    size_t head = (16 - (a & 15)) & 15;
    if (head > bytes) head = bytes;
    size_t nvec = (bytes - head) / 16, tail = bytes - head - nvec * 16;
    unsigned int b = (unsigned int) (value & 0xff);
    unsigned int word = b | (b << 8) | (b << 16) | (b << 24);
    size_t work = nvec > head + tail ? nvec : head + tail;
    size_t blocks = (work + 255) / 256;
    size_t cap = (size_t) ca3::cu_count() * 16;
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
    fill_bytes_kernel<<<(unsigned) blocks, 256, 0, ca3::stream()>>>((unsigned char *) dptr, head, nvec,
                                                                     tail, word);
    return ca3::finish();
}

int hip_free_dev(void *dptr)
{
    if (!dptr) return HIP_PROXY_OK;
    CA3_CHECK(hipFree(dptr));
    return HIP_PROXY_OK;
}

int hip_free_host(void *hptr)
{
    if (!hptr) return HIP_PROXY_OK;
    CA3_CHECK(hipHostFree(hptr));
    return HIP_PROXY_OK;
}

static int copy_kind(const void *src, void *dst, const size_t bytes, hipMemcpyKind kind)
{
    if (bytes == 0) return HIP_PROXY_OK;
    if (!src || !dst) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "memcpy: null pointer");
    CA3_CHECK(hipMemcpyAsync(dst, src, bytes, kind, ca3::stream()));
    // Host-side copies always complete before returning: the caller may
    // reuse or free the host buffer right away.
    if (kind == hipMemcpyDeviceToDevice) return ca3::finish();
    CA3_CHECK(hipStreamSynchronize(ca3::stream()));
    return HIP_PROXY_OK;
}

int hip_memcpy_h2d(const void *hptr, void *dptr, const size_t bytes)
{ return copy_kind(hptr, dptr, bytes, hipMemcpyHostToDevice); }

int hip_memcpy_d2h(const void *dptr, void *hptr, const size_t bytes)
{ return copy_kind(dptr, hptr, bytes, hipMemcpyDeviceToHost); }

int hip_memcpy_d2d(const void *dptr_src, void *dptr_dst, const size_t bytes)
{ return copy_kind(dptr_src, dptr_dst, bytes, hipMemcpyDeviceToDevice); }

int hip_memcpy_auto(const void *src, void *dst, const size_t bytes)
{
    if (bytes == 0) return HIP_PROXY_OK;
    CA3_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ca3::stream()));
    CA3_CHECK(hipStreamSynchronize(ca3::stream()));
    return HIP_PROXY_OK;
}

}  // extern "C"
