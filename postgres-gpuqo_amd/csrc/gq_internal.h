/* gq_internal.h -- library-internal helpers shared by the host sources. */
#ifndef GQ_INTERNAL_H
#define GQ_INTERNAL_H

#include "../../include/gpuqo_amd.h"

namespace gq {

/* Record the calling thread's last error message (printf format). */
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

/* Bind the device context (stream, constant tables) to `device`. */
int device_init(int device);

/* Interrupt flag polled between DP calls (CHECK_FOR_INTERRUPTS). */
bool interrupted();

}  // namespace gq

#endif
