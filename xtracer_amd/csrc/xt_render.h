// Shared pieces of the render loop used by one-job renders and the server.
#pragma once
#include "xt_device.h"

int xt_job_check(xt_dev *d, const xt_job *job, JobDev &j, const DevScene **ds_out);
uint32_t xt_rounds(const JobDev &j);
uint32_t xt_auto_spp_pass(const JobDev &j, uint32_t requested, uint32_t target_paths);
int xt_launch_shade(xt_dev *d, const DevScene &ds, const JobDev &j, uint32_t r, int qi, size_t cap);
