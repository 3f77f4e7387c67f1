// Host-only stand-in for the few HIP runtime types and calls the serving
// loop uses, so its threading can be exercised on the CPU
// (tests/native/serve_sim.cpp supplies the function bodies).
#pragma once
#include <stddef.h>
#include <stdint.h>
typedef int hipError_t;
typedef struct ihipStream_t *hipStream_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2, hipErrorNotInitialized = 3 };
enum hipMemcpyKind { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2 };
hipError_t hipGetDeviceCount(int *n);
hipError_t hipMalloc(void **p, size_t bytes);
template <typename T>
hipError_t hipMalloc(T **p, size_t bytes) { return hipMalloc((void **)p, bytes); }
hipError_t hipFree(void *p);
hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind k, hipStream_t s);
hipError_t hipStreamCreate(hipStream_t *s);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
enum { hipStreamNonBlocking = 1 };
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags);
hipError_t hipGetDevice(int *d);
// stream-ordered memory pool
typedef struct ihipMemPool_t *hipMemPool_t;
enum hipMemAllocationType { hipMemAllocationTypePinned = 1 };
enum hipMemLocationType { hipMemLocationTypeDevice = 1 };
struct hipMemLocation { hipMemLocationType type; int id; };
struct hipMemPoolProps { hipMemAllocationType allocType; int handleTypes; hipMemLocation location; };
enum hipMemPoolAttr { hipMemPoolAttrReleaseThreshold = 4, hipMemPoolAttrReservedMemHigh = 6,
                      hipMemPoolAttrUsedMemHigh = 8 };
hipError_t hipMemPoolCreate(hipMemPool_t *pool, const hipMemPoolProps *props);
hipError_t hipMemPoolDestroy(hipMemPool_t pool);
hipError_t hipMemPoolSetAttribute(hipMemPool_t pool, hipMemPoolAttr attr, void *value);
hipError_t hipMemPoolGetAttribute(hipMemPool_t pool, hipMemPoolAttr attr, void *value);
hipError_t hipMallocFromPoolAsync(void **p, size_t bytes, hipMemPool_t pool, hipStream_t s);
hipError_t hipFreeAsync(void *p, hipStream_t s);
