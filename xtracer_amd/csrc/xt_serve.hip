// Continuous-batching render server (xt_submit / xt_poll / xt_serve_step).
#include <hip/hip_runtime.h>

#include "xt_device.h"

void xt_serve_shutdown(xt_dev *d) { (void)d; }

extern "C" int xt_submit(xt_dev *d, const xt_job *job, uint64_t *ticket) {
    (void)job;
    (void)ticket;
    if (d) d->err = "serving not available";
    return XT_E_STATE;
}

extern "C" int xt_poll(xt_dev *d, uint64_t ticket, int *done, float *rgba_out, xt_job_stats *stats) {
    (void)ticket;
    (void)done;
    (void)rgba_out;
    (void)stats;
    if (d) d->err = "serving not available";
    return XT_E_STATE;
}

extern "C" int xt_serve_step(xt_dev *d, int *active_jobs) {
    (void)active_jobs;
    if (d) d->err = "serving not available";
    return XT_E_STATE;
}

extern "C" int xt_drain(xt_dev *d) {
    if (d) d->err = "serving not available";
    return XT_E_STATE;
}
