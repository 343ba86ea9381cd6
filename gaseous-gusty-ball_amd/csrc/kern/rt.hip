// rt.hip -- device runtime behind the thin C ABI: one device, one default-priority
// stream, device/pinned allocation, synchronous copies, graph-safe zero fill, events,
// and precision conversion kernels.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include "dev.h"
#include "ggbk.h"

static hipStream_t g_stream = nullptr;
static int g_ready = 0;
static char g_err[512];

static int fail(const char *where, hipError_t e)
{
	snprintf(g_err, sizeof(g_err), "%s: %s", where, hipGetErrorString(e));
	fprintf(stderr, "ggb: %s\n", g_err);
	return -1;
}

hipStream_t ggbk_stream(void) { return g_stream; }

extern "C" void *ggbk_stream_handle(void) { return (void *) g_stream; }

extern "C" const char *ggbk_last_error(void) { return g_err; }

extern "C" int ggbk_init(void)
{
	if (g_ready)
		return 0;
	int ndev = 0;
	hipError_t e = hipGetDeviceCount(&ndev);
	if (e != hipSuccess || ndev < 1)
		return fail("ggbk_init: no HIP device", e == hipSuccess ? hipErrorNoDevice : e);
	e = hipSetDevice(0);
	if (e != hipSuccess)
		return fail("ggbk_init: hipSetDevice", e);
	e = hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking);
	if (e != hipSuccess)
		return fail("ggbk_init: hipStreamCreate", e);
	if (ggbk_noalloc_init()) {
		snprintf(g_err, sizeof(g_err), "ggbk_init: the device scratch could not be located");
		fprintf(stderr, "ggb: %s\n", g_err);
		return -1;
	}
	g_ready = 1;
	return 0;
}

extern "C" void *ggbk_malloc(size_t bytes)
{
	void *p = nullptr;
	if (bytes == 0)
		bytes = 256;
	hipError_t e = hipMalloc(&p, bytes);
	if (e != hipSuccess) {
		fail("hipMalloc", e);
		return nullptr;
	}
	return p;
}

extern "C" void ggbk_free(void *p)
{
	if (p)
		(void) hipFree(p);
}

extern "C" void *ggbk_host_alloc(size_t bytes)
{
	void *p = nullptr;
	if (hipHostMalloc(&p, bytes ? bytes : 64, hipHostMallocDefault) != hipSuccess)
		return nullptr;
	return p;
}

extern "C" void ggbk_host_free(void *p)
{
	if (p)
		(void) hipHostFree(p);
}

extern "C" int ggbk_h2d(void *dst, const void *src, size_t bytes)
{
	if (!bytes)
		return 0;
	hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_stream);
	if (e == hipSuccess)
		e = hipStreamSynchronize(g_stream);
	return e == hipSuccess ? 0 : fail("h2d", e);
}

extern "C" int ggbk_d2h(void *dst, const void *src, size_t bytes)
{
	if (!bytes)
		return 0;
	hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_stream);
	if (e == hipSuccess)
		e = hipStreamSynchronize(g_stream);
	return e == hipSuccess ? 0 : fail("d2h", e);
}

extern "C" int ggbk_d2h_async(void *dst, const void *src, size_t bytes)
{
	if (!bytes)
		return 0;
	hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_stream);
	return e == hipSuccess ? 0 : fail("d2h_async", e);
}

extern "C" int ggbk_d2d(void *dst, const void *src, size_t bytes)
{
	if (!bytes)
		return 0;
	hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, g_stream);
	return e == hipSuccess ? 0 : fail("d2d", e);
}

__global__ void k_zero_words(uint32_t *p, size_t nwords)
{
	size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
	size_t stride = (size_t) gridDim.x * blockDim.x;
	for (; i < nwords; i += stride)
		p[i] = 0u;
}

__global__ void k_zero_bytes(unsigned char *p, size_t nbytes)
{
	size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
	if (i < nbytes)
		p[i] = 0;
}

extern "C" int ggbk_zero(void *p, size_t bytes)
{
	if (!bytes)
		return 0;
	size_t words = bytes / 4, tail = bytes % 4;
	if (words) {
		long long blocks = (long long) ((words + 255) / 256);
		if (blocks > 65536)
			blocks = 65536;
		GGBK_LAUNCH(k_zero_words, (unsigned) blocks, 256, 0, (uint32_t *) p, words);
	}
	if (tail)
		GGBK_LAUNCH(k_zero_bytes, 1, 64, 0, (unsigned char *) p + 4 * words, tail);
	return ggbk_check("zero");
}

extern "C" int ggbk_sync(void)
{
	hipError_t e = hipStreamSynchronize(g_stream);
	return e == hipSuccess ? 0 : fail("sync", e);
}

extern "C" int ggbk_check(const char *where)
{
	hipError_t e = hipGetLastError();
	return e == hipSuccess ? 0 : fail(where, e);
}

extern "C" void *ggbk_event_new(void)
{
	hipEvent_t ev;
	if (hipEventCreate(&ev) != hipSuccess)
		return nullptr;
	return (void *) ev;
}

extern "C" void ggbk_event_record(void *ev) { (void) hipEventRecord((hipEvent_t) ev, g_stream); }

extern "C" float ggbk_event_ms(void *a, void *b)
{
	float ms = 0;
	(void) hipEventSynchronize((hipEvent_t) b);
	(void) hipEventElapsedTime(&ms, (hipEvent_t) a, (hipEvent_t) b);
	return ms;
}

extern "C" void ggbk_event_free(void *ev)
{
	if (ev)
		(void) hipEventDestroy((hipEvent_t) ev);
}

// ---------------- launch probes ----------------
// Selected launchers bracket their kernel with HIP events on the library stream when
// the probe of their id is enabled; the measurement tools sum the intervals.
#define PROBE_IDS 16
#define PROBE_CAP 8192
static unsigned g_probe_mask = 0;
static hipEvent_t g_probe_ev[PROBE_IDS][PROBE_CAP][2];
static int g_probe_n[PROBE_IDS];
static int g_probe_open[PROBE_IDS];

extern "C" void ggbk_probe_enable(int id, int on)
{
	if (id < 0 || id >= PROBE_IDS)
		return;
	if (on)
		g_probe_mask |= 1u << id;
	else
		g_probe_mask &= ~(1u << id);
}

int ggbk_probe_begin(int id)
{
	if (!(g_probe_mask & (1u << id)) || g_probe_n[id] >= PROBE_CAP)
		return 0;
	hipEvent_t *e = g_probe_ev[id][g_probe_n[id]];
	if (!e[0]) {
		(void) hipEventCreate(&e[0]);
		(void) hipEventCreate(&e[1]);
	}
	(void) hipEventRecord(e[0], g_stream);
	g_probe_open[id] = 1;
	return 1;
}

void ggbk_probe_end(int id)
{
	if (!g_probe_open[id])
		return;
	(void) hipEventRecord(g_probe_ev[id][g_probe_n[id]][1], g_stream);
	g_probe_n[id]++;
	g_probe_open[id] = 0;
}

// Number of probed launches and their summed duration (ms) since the last collect;
// synchronises and resets the probe.
extern "C" int ggbk_probe_collect(int id, double *total_ms)
{
	*total_ms = 0;
	if (id < 0 || id >= PROBE_IDS)
		return 0;
	(void) hipStreamSynchronize(g_stream);
	int n = g_probe_n[id];
	for (int k = 0; k < n; k++) {
		float ms = 0;
		(void) hipEventElapsedTime(&ms, g_probe_ev[id][k][0], g_probe_ev[id][k][1]);
		*total_ms += ms;
	}
	g_probe_n[id] = 0;
	return n;
}

// A plain copy of `bytes` bytes (float words) for counter calibration.
extern "C" void ggbk_copy_kernel(void *dst, const void *src, size_t bytes)
{
	ggbk_convert(dst, GGBK_F32, src, GGBK_F32, bytes / 4);
}

template <typename D, typename S>
__global__ void k_convert(D *dst, const S *src, size_t count)
{
	size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
	size_t stride = (size_t) gridDim.x * blockDim.x;
	for (; i < count; i += stride)
		dst[i] = (D) src[i];
}

extern "C" void ggbk_convert(void *dst, int tdst, const void *src, int tsrc, size_t count)
{
	if (!count)
		return;
	long long blocks = (long long) ((count + 255) / 256);
	if (blocks > 65536)
		blocks = 65536;
	unsigned nb = (unsigned) blocks;
	if (tdst == GGBK_F64 && tsrc == GGBK_F64)
		GGBK_LAUNCH((k_convert<double, double>), nb, 256, 0, (double *) dst, (const double *) src, count);
	else if (tdst == GGBK_F64)
		GGBK_LAUNCH((k_convert<double, float>), nb, 256, 0, (double *) dst, (const float *) src, count);
	else if (tsrc == GGBK_F64)
		GGBK_LAUNCH((k_convert<float, double>), nb, 256, 0, (float *) dst, (const double *) src, count);
	else
		GGBK_LAUNCH((k_convert<float, float>), nb, 256, 0, (float *) dst, (const float *) src, count);
}
