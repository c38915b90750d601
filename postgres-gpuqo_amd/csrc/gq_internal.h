/* gq_internal.h -- library-internal helpers shared by the host sources. */
#ifndef GQ_INTERNAL_H
#define GQ_INTERNAL_H

#include <exception>
#include <new>

#include "../../include/gpuqo_amd.h"

namespace gq {

/* Record the calling thread's last error message (printf format). */
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

/* Check a query's sizes, arrays, widths and class members (GQ_OK or
 * GQ_ERR_INVALID with the message set). */
int validate(const gq_query *q);

/* Check the config fields with a closed set of values the DP reads
 * (cost_function), GQ_OK or GQ_ERR_INVALID with the message set. */
int validate_config(const gq_config *c);

/* Bind the device context (stream, constant tables) to `device`. */
int device_init(int device);

/* The device context is bound to `device` (gq_init has run for it). */
bool device_ready(int device);

/* Interrupt flag (CHECK_FOR_INTERRUPTS), polled between DP calls and
 * between the level launches of one DP. */
bool interrupted();

/* Body of an exported entry point: no C++ exception leaves the library.
 * std::bad_alloc becomes GQ_ERR_NOMEM, anything else GQ_ERR_INTERNAL
 * (e.g. std::system_error when a host thread cannot be started). */
template <class F>
gq_status guarded(F &&body) noexcept {
    try {
        return (gq_status)body();
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return GQ_ERR_NOMEM;
    } catch (const std::exception &e) {
        set_error("internal error: %s", e.what());
        return GQ_ERR_INTERNAL;
    } catch (...) {
        set_error("internal error");
        return GQ_ERR_INTERNAL;
    }
}

}  // namespace gq

#endif
