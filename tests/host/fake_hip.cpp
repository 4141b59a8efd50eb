// A stand-in HIP runtime for checking the library's HOST code on a machine
// without a GPU: the planner, the C ABI's staging and argument checks, and
// the serving loop's admission, retirement and error paths, compiled for the
// host only and run under AddressSanitizer / UndefinedBehaviorSanitizer.
//
// Device memory is zero-filled host memory, copies are memcpy, kernel
// launches are counted and do nothing, events and streams complete at once.  So the
// numbers the library computes are meaningless here; what is checked is
// that its host code never reads or writes out of bounds, leaks, or
// mishandles a failing allocation.  fake_hip_fail_after(k) makes every
// device allocation after the next k fail, -1 turns that off.
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace {
long g_fail_after = -1;  // allocations left before they start failing (-1: never)
long g_launches = 0;
long g_live_allocs = 0;
dim3 g_grid, g_block;
size_t g_shmem = 0;
hipStream_t g_stream = nullptr;

struct Obj {
  int kind;
};

bool alloc_ok() {
  if (g_fail_after < 0) return true;
  if (g_fail_after == 0) return false;
  --g_fail_after;
  return true;
}
}  // namespace

extern "C" {

void fake_hip_fail_after(long k) { g_fail_after = k; }
long fake_hip_launches() { return g_launches; }
long fake_hip_live_allocs() { return g_live_allocs; }

// --- registration of the (absent) device code ---
void** __hipRegisterFatBinary(const void*) {
  static void* handle = nullptr;
  return &handle;
}
void __hipRegisterFunction(void**, const void*, char*, const char*, unsigned int, uint3*, uint3*, dim3*, dim3*, int*) {}
void __hipRegisterVar(void**, void*, char*, char*, int, size_t, int, int) {}
void __hipUnregisterFatBinary(void**) {}

hipError_t __hipPushCallConfiguration(dim3 gridDim, dim3 blockDim, size_t sharedMem, hipStream_t stream) {
  g_grid = gridDim;
  g_block = blockDim;
  g_shmem = sharedMem;
  g_stream = stream;
  return hipSuccess;
}
hipError_t __hipPopCallConfiguration(dim3* gridDim, dim3* blockDim, size_t* sharedMem, hipStream_t* stream) {
  *gridDim = g_grid;
  *blockDim = g_block;
  *sharedMem = g_shmem;
  *stream = g_stream;
  return hipSuccess;
}
hipError_t hipLaunchKernel(const void*, dim3 numBlocks, dim3 dimBlocks, void**, size_t, hipStream_t) {
  if (numBlocks.x == 0 || dimBlocks.x == 0 || dimBlocks.x > 1024) {
    fprintf(stderr, "fake_hip: bad launch grid %u block %u\n", numBlocks.x, dimBlocks.x);
    abort();
  }
  ++g_launches;
  return hipSuccess;
}

// --- device and errors ---
hipError_t hipGetDeviceCount(int* count) {
  *count = 1;
  return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipDeviceGetAttribute(int* pi, hipDeviceAttribute_t attr, int) {
  switch (attr) {
    case hipDeviceAttributeMultiprocessorCount: *pi = 256; break;
    case hipDeviceAttributeMaxSharedMemoryPerBlock: *pi = 160 * 1024; break;
    case hipDeviceAttributeWallClockRate: *pi = 100000; break;
    default: *pi = 1; break;
  }
  return hipSuccess;
}
hipError_t hipGetLastError(void) { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "fake_hip error"; }
hipError_t hipFuncSetAttribute(const void*, hipFuncAttribute, int) { return hipSuccess; }
hipError_t hipGetSymbolAddress(void** devPtr, const void* symbol) {
  *devPtr = const_cast<void*>(symbol);  // the host shadow stands in for the device variable
  return hipSuccess;
}

// --- memory ---
hipError_t hipMalloc(void** ptr, size_t size) {
  *ptr = nullptr;
  if (!alloc_ok()) return hipErrorOutOfMemory;
  const size_t bytes = (size + 255) & ~(size_t)255;
  *ptr = aligned_alloc(256, bytes);
  if (!*ptr) return hipErrorOutOfMemory;
  memset(*ptr, 0, bytes);  // kernels do not run here: status words read back as 0
  ++g_live_allocs;
  return hipSuccess;
}
hipError_t hipFree(void* ptr) {
  if (ptr) {
    free(ptr);
    --g_live_allocs;
  }
  return hipSuccess;
}
hipError_t hipHostMalloc(void** ptr, size_t size, unsigned int) { return hipMalloc(ptr, size); }
hipError_t hipHostFree(void* ptr) { return hipFree(ptr); }
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind) {
  if (bytes) memmove(dst, src, bytes);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind k, hipStream_t) {
  return hipMemcpy(dst, src, bytes, k);
}
hipError_t hipMemcpy2DAsync(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width, size_t height,
                            hipMemcpyKind, hipStream_t) {
  for (size_t r = 0; r < height; ++r) memmove((char*)dst + r * dpitch, (const char*)src + r * spitch, width);
  return hipSuccess;
}
hipError_t hipMemset2DAsync(void* dst, size_t pitch, int value, size_t width, size_t height, hipStream_t) {
  for (size_t r = 0; r < height; ++r) memset((char*)dst + r * pitch, value, width);
  return hipSuccess;
}

int hip_malloc_for_tests(void** ptr, size_t bytes) { return hipMalloc(ptr, bytes) == hipSuccess ? 0 : 1; }
int hip_free_for_tests(void* ptr) { return hipFree(ptr) == hipSuccess ? 0 : 1; }

// --- streams, events, graphs: everything has completed ---
hipError_t hipStreamCreate(hipStream_t* s) {
  *s = reinterpret_cast<hipStream_t>(new Obj{1});
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  delete reinterpret_cast<Obj*>(s);
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned int) { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t* e) {
  *e = reinterpret_cast<hipEvent_t>(new Obj{2});
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) {
  delete reinterpret_cast<Obj*>(e);
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) {
  *ms = 0.0f;
  return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipSuccess; }
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
  *g = reinterpret_cast<hipGraph_t>(new Obj{3});
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* x, hipGraph_t, hipGraphNode_t*, char*, size_t) {
  *x = reinterpret_cast<hipGraphExec_t>(new Obj{4});
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipSuccess; }
hipError_t hipGraphDestroy(hipGraph_t g) {
  delete reinterpret_cast<Obj*>(g);
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t x) {
  delete reinterpret_cast<Obj*>(x);
  return hipSuccess;
}

}  // extern "C"
