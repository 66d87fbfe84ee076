/*
 * Device copy of known size, for calibrating the HBM byte counters of the
 * profiler (FETCH_SIZE / WRITE_SIZE) and for a measured copy bandwidth.
 */
#include "ctx.h"

__global__ __launch_bounds__(256) void k_copy_calib(const double2 *__restrict__ src,
						    double2 *__restrict__ dst, long n)
{
	const long stride = (long)gridDim.x * blockDim.x;
	for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
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
	const int grid = ncu * 16;
	hipEvent_t e0, e1;
	HIPCHECK(hipEventCreate(&e0));
	HIPCHECK(hipEventCreate(&e1));
	k_copy_calib<<<grid, 256>>>(a, b, n); /* warm-up */
	HIPCHECK(hipEventRecord(e0, 0));
	for (int r = 0; r < reps; r++)
		k_copy_calib<<<grid, 256>>>(a, b, n);
	HIPCHECK(hipEventRecord(e1, 0));
	HIPCHECK(hipEventSynchronize(e1));
	float ms = 0.f;
	HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
	HIPCHECK(hipEventDestroy(e0));
	HIPCHECK(hipEventDestroy(e1));
	HIPCHECK(hipFree(a));
	HIPCHECK(hipFree(b));
	return (double)ms / reps;
}
