/* Wall-clock timers (src/utils/timing.c:16-29). */
#include <stddef.h>
#include <utils/timing.h>

void tic(struct timeval *start)
{
	gettimeofday(start, NULL);
}

gdn_real toc(struct timeval *start)
{
	struct timeval end;
	gettimeofday(&end, NULL);
	return (gdn_real)(end.tv_sec - start->tv_sec) +
	       (gdn_real)(end.tv_usec - start->tv_usec) * 1e-6;
}
