/* Wall-clock timers. Replaces src/utils/timing.h (src/utils/timing.c:16-29). */
#ifndef UTILS_TIMING_H
#define UTILS_TIMING_H

#include <sys/time.h>
#include <gdon3d.h>

#ifdef __cplusplus
extern "C" {
#endif

void tic(struct timeval *start);
gdn_real toc(struct timeval *start);

#ifdef __cplusplus
}
#endif
#endif
