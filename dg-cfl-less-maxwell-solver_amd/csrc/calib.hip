/*
 * Device copy of known size, for calibrating the HBM byte counters of the
 * profiler (FETCH_SIZE / WRITE_SIZE) and for a measured copy bandwidth.
 */
#include "ctx.h"

__global__ __launch_bounds__(256) void k_copy_calib(const double2 *__restrict__ src,
						    double2 *__restrict__ dst, long n)
{
	/* tiles of 4 x 256 consecutive elements (16 KB), taken grid-stride: each
	 * thread has four independent 16-byte loads in flight before their
	 * stores, and each block's loads stay within one contiguous tile */
	const long ntile = n / 1024;
	for (long tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
		const long k = tile * 1024 + threadIdx.x;
		double2 a[4];
#pragma unroll
		for (int u = 0; u < 4; u++)
			a[u] = src[k + 256 * u];
#pragma unroll
		for (int u = 0; u < 4; u++)
			dst[k + 256 * u] = a[u];
	}
	for (long k = ntile * 1024 + (long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
	     k += (long)gridDim.x * blockDim.x)
		dst[k] = src[k];
}

extern "C" double gdh_copy_calibration(size_t bytes, int reps)
{
	const long n = (long)(bytes / sizeof(double2));
	if (n <= 0 || reps <= 0)
		return 0.0;
	double2 *a = (double2 *)dmalloc(sizeof(double2) * (size_t)n);
	double2 *b = (double2 *)dmalloc(sizeof(double2) * (size_t)n);
	HIPCHECK(hipMemset(a, 0, sizeof(double2) * (size_t)n));
	int dev = 0, ncu = 0;
	HIPCHECK(hipGetDevice(&dev));
	HIPCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
	hipEvent_t e0, e1;
	HIPCHECK(hipEventCreate(&e0));
	HIPCHECK(hipEventCreate(&e1));
	/* 8, 16, 32 or 64 blocks of 256 threads per CU: the fastest mean */
	double best = 0.0;
	for (int per_cu = 8; per_cu <= 64; per_cu *= 2) {
		const int grid = ncu * per_cu;
		k_copy_calib<<<grid, 256>>>(a, b, n); /* warm-up */
		HIPCHECK(hipEventRecord(e0, 0));
		for (int r = 0; r < reps; r++)
			k_copy_calib<<<grid, 256>>>(a, b, n);
		HIPCHECK(hipEventRecord(e1, 0));
		HIPCHECK(hipEventSynchronize(e1));
		float ms = 0.f;
		HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
		if (best == 0.0 || ms / reps < best)
			best = (double)ms / reps;
	}
	HIPCHECK(hipEventDestroy(e0));
	HIPCHECK(hipEventDestroy(e1));
	HIPCHECK(hipFree(a));
	HIPCHECK(hipFree(b));
	return best;
}
